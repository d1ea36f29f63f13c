"""Histogram files (host code in libwdl.so, no GPU): the CSV / RAW exporters and
the 1-D CSV loader of histogramIO.h.  Pinned against what the formats fix: RAW
is the doubles' bytes (numpy.tofile), CSV is C++ stream output at its default
precision 6 (printf %g), and the loader's normalisation and its read loop
(histogramIO.h:323-328: a trailing newline repeats the last value, a file
without one does not), checked on hand-written files with known answers."""
import os

import numpy as np

import wassersteindictionarylearning_amd as wdl


def test_raw_is_native_doubles(tmp_path):
    v = np.random.default_rng(3).random(257) * 1e-3
    assert wdl.export_histogram(tmp_path, "h0", v, "raw") == 0
    assert (tmp_path / "h0.raw").read_bytes() == v.tobytes()


def test_csv_default_stream_precision(tmp_path):
    v = np.array([0.123456789, 1e-7, 2.5, 1234567.0, 0.0, 1.0 / 3.0])
    assert wdl.export_histogram(tmp_path, "h1", v) == 0
    lines = (tmp_path / "h1.csv").read_text().split("\n")
    assert lines == ["0.123457", "1e-07", "2.5", "1.23457e+06", "0", "0.333333", ""]
    assert lines[:-1] == ["%g" % x for x in v]


def test_batch_and_unwritable_folder(tmp_path):
    hs = [np.arange(3.0), np.ones(5)]
    assert wdl.export_histograms(tmp_path, ["a", "b"], hs, "raw") == 0
    assert np.array_equal(np.fromfile(tmp_path / "b.raw"), hs[1])
    # a missing folder: every file fails, -1 as the reference
    assert wdl.export_histograms(tmp_path / "missing", ["a", "b"], hs) == -1
    assert "Unable to write file" in wdl.lib().wdl_last_error().decode()


def test_load_csv_known_answers(tmp_path):
    f = tmp_path / "x.csv"
    f.write_text("1\n2\n5")  # no trailing newline: three values, sum 8
    s, v = wdl.load_csv_to_pdf(f)
    assert s == 8.0 and np.array_equal(v, np.array([1, 2, 5]) / 8.0)
    f.write_text("1\n2\n5\n")  # trailing newline: the failed read repeats 5
    s, v = wdl.load_csv_to_pdf(f)
    assert s == 13.0 and np.array_equal(v, np.array([1, 2, 5, 5]) / 13.0)
    s, v = wdl.load_csv_to_pdf(tmp_path / "absent.csv")
    assert s == -1.0 and v.size == 0


def test_csv_round_trip(tmp_path):
    v = np.random.default_rng(5).random(100)
    wdl.export_histogram(tmp_path, "r", v)
    s, back = wdl.load_csv_to_pdf(tmp_path / "r.csv")
    # six digits written; the reloaded copy carries the last value twice
    assert back.size == 101
    ref = np.array([float("%g" % x) for x in v] + [float("%g" % v[-1])])
    assert abs(s - ref.sum()) < 1e-12 * ref.sum()
    assert np.allclose(back, ref / ref.sum(), rtol=1e-15, atol=0)
