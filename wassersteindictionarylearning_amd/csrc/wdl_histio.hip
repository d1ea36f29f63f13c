// Histogram files on either side of the solver: the CSV and RAW exporters
// (ExportHistogramToCSV / ExportHistogramToRAW, histogramIO.h:56-135, 202-240)
// and the 1-D CSV loader (load_csv_to_pdf, histogramIO.h:314-337).  Host code
// only; these are the formats the reference's dictionaries, weights and
// barycenters are written in and its 1-D inputs are read from.
#include <fstream>
#include <string>
#include <vector>

#include "wdl_internal.h"
#include "../../include/wdl_c.h"

namespace {

std::string join_path(const char* folder, const char* name, const char* ext) {
  return std::string(folder) + "/" + name + ext;
}

// One file per histogram; -1 (and the message) when it cannot be opened, as the
// reference's exportHistogram (histogramIO.h:62-78, 208-220).
int export_one(const char* folder, const char* name, const double* v, long long n, bool raw) {
  const std::string fn = join_path(folder, name, raw ? ".raw" : ".csv");
  std::ofstream ofs(fn, raw ? std::ios::out | std::ios::binary : std::ios::out);
  if (!ofs.is_open()) {
    wdl::set_error("Unable to write file " + fn);
    return -1;
  }
  if (raw) {
    ofs.write(reinterpret_cast<const char*>(v), n * (long long)sizeof(double));
  } else {
    // default stream formatting (6 significant digits), one value per line
    for (long long i = 0; i < n; ++i) ofs << v[i] << "\n";
  }
  return 0;
}

int export_batch(const char* folder, const char* const* names, const double* const* values,
                 const long long* lengths, int count, bool raw) {
  if (!folder || count < 0 || (count > 0 && (!names || !values || !lengths))) {
    wdl::set_error("export: null argument or negative count");
    return WDL_ERR_ARG;
  }
  // every file is attempted; -1 if any failed (histogramIO.h:82-97)
  bool ok = true;
  for (int i = 0; i < count; ++i) {
    if (!names[i] || lengths[i] < 0 || (lengths[i] > 0 && !values[i])) {
      wdl::set_error("export: null name or values, or negative length");
      return WDL_ERR_ARG;
    }
    if (export_one(folder, names[i], values[i], lengths[i], raw) != 0) ok = false;
  }
  return ok ? 0 : -1;
}

}  // namespace

extern "C" {

int wdl_export_histogram_csv(const char* folder, const char* name, const double* values,
                             long long n) {
  return export_batch(folder, &name, &values, &n, 1, false);
}

int wdl_export_histogram_raw(const char* folder, const char* name, const double* values,
                             long long n) {
  return export_batch(folder, &name, &values, &n, 1, true);
}

int wdl_export_histograms_csv(const char* folder, const char* const* names,
                              const double* const* values, const long long* lengths, int count) {
  return export_batch(folder, names, values, lengths, count, false);
}

int wdl_export_histograms_raw(const char* folder, const char* const* names,
                              const double* const* values, const long long* lengths, int count) {
  return export_batch(folder, names, values, lengths, count, true);
}

double wdl_load_csv_to_pdf(const char* filename, double* out, long long capacity,
                           long long* n_out) {
  if (!filename || !n_out || capacity < 0 || (capacity > 0 && !out)) {
    wdl::set_error("load_csv_to_pdf: null argument or negative capacity");
    return -1.0;
  }
  std::ifstream ifs(filename);
  if (!ifs.is_open()) {
    wdl::set_error(std::string("Unable to read file ") + filename);
    return -1.0;
  }
  // The reference's loop (histogramIO.h:323-328) tests the stream before each
  // read and keeps the value of a read that failed: a file that ends in
  // whitespace yields its last value once more (0 for an empty file), and that
  // entry counts in the sum.  The same stream reads reproduce it here.
  std::vector<double> v;
  double x = 0.0, sum = 0.0;
  while (ifs.good()) {
    ifs >> x;
    sum += x;
    v.push_back(x);
  }
  *n_out = (long long)v.size();
  for (long long i = 0; i < (long long)v.size() && i < capacity; ++i) out[i] = v[i] / sum;
  return sum;
}

}  // extern "C"
