"""The C-ABI library loads and exports every entry point include/wdl_c.h declares (CPU only)."""
import ctypes
import os
import re

import pytest

import wassersteindictionarylearning_amd as wdl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wdl_c.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wdl_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_header():
    assert os.path.exists(wdl.LIB_PATH), "libwdl.so not built"
    lib = ctypes.CDLL(wdl.LIB_PATH)
    names = declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python layer binds exactly the declared surface
    assert sorted(wdl._SIGS) == names


def test_version_and_footprint():
    assert wdl.lib().wdl_version().decode().startswith("wdl-gfx950")
    W, H, K, P, L = 128, 128, 5, 64, 100
    f = wdl.SinkhornSolver.footprint(W, H, K, P, L)
    PKN = P * K * W * H
    hist = 2 * L * PKN * 4 + L * P * W * H * 4   # conv_b, phi (FP32) and q per step
    assert hist <= f < hist * 1.1
    logf = wdl.SinkhornSolver.footprint(W, H, K, P, L, wdl.FLAG_LOG_DOMAIN)
    assert logf >= 3 * L * PKN * 8               # log b, conv_b, phi in FP64
    assert wdl.SinkhornSolver.footprint(W, H, K, P, L, wdl.FLAG_WARM_RESTART) == f + 2 * PKN * 4
    assert wdl.lib().wdl_solver_footprint(0, H, K, P, L, 0) == -1
    # tiled sizes add the second-convolution inputs, per-tile dots and log-phi sums
    ft = wdl.SinkhornSolver.footprint(W, H, K, P, L, wdl.FLAG_TILED)
    assert ft == f + 2 * PKN * 4 + P * K * 4 * 8 * (1 + L)
    ftl = wdl.SinkhornSolver.footprint(W, H, K, P, L, wdl.FLAG_TILED | wdl.FLAG_LOG_DOMAIN)
    # log domain: also the backward's cross-atom sums (two P x N FP32 images)
    assert ftl == logf + 2 * PKN * 9 + P * K * 4 * 8 * (1 + L) + 2 * P * W * H * 4


def test_null_arguments_are_errors():
    L = wdl.lib()
    assert L.wdl_gauss2d_convolve(None, None, None, 1) == -1
    assert b"NULL" in L.wdl_last_error()
    assert L.wdl_solver_run(None, 0) == -1
    assert L.wdl_ctx_create(0, None) == -1
    # serving entry points check their arguments before touching the device
    assert L.wdl_solver_set_slot(None, 0, None, None, 0, 0.0) == -1
    assert L.wdl_solver_projection_round(None) == -1
    assert L.wdl_solver_get_slot(None, 0, None, None) == -1
    assert L.wdl_ctx_query(None) == -1
    assert L.wdl_server_create(0, 1 << 30, 4, None) == -1
    assert L.wdl_server_create(0, 0, 4, None) == -1
    assert L.wdl_server_submit(None, 0, None, 1, 1, 0.5, None) == -1
    assert L.wdl_server_wait(None, 1, 0.0, None, None, None, None) == -1
    assert L.wdl_server_stats(None, None) == -1
    assert L.wdl_server_destroy(None) == 0


def test_checkpoint_footprint():
    # the 256x256 log workload (K=6, P=32, L=200): 61 GB of histories in full;
    # checkpointed: a ring of ceil(sqrt(200)) + 2 = 17 slots and 13 checkpoints
    LOG, CK = wdl.FLAG_LOG_DOMAIN, wdl.FLAG_CHECKPOINT
    full = wdl.SinkhornSolver.footprint(256, 256, 6, 32, 200, LOG)
    ck = wdl.SinkhornSolver.footprint(256, 256, 6, 32, 200, LOG | CK)
    PKN = 32 * 6 * 256 * 256
    assert full - ck == (3 * (200 - 17) - 13) * PKN * 8
    assert ck < 0.2 * full
    # too few iterations for a ring smaller than L: the full histories
    assert wdl.SinkhornSolver.footprint(64, 64, 3, 4, 6, CK) == \
        wdl.SinkhornSolver.footprint(64, 64, 3, 4, 6, 0)
