// gahhg_cli.cpp -- command-line driver with the reference's flags
// (source/GPISP.cu:15-70):
//   gahhg /polyfile P /tpfile Q /reportfile R /resultfile S [/maxdepth D]
// Defaults as GPISPParam (source/GPISPParam.h:66-77).  Unknown flags are
// skipped like the reference's loop does.  /maxdepth D (the AHHG.exe flag,
// test/testAHHG.bat) builds exactly D levels with the density stop rule
// off; without it the auto-depth rule runs with at most 15 levels.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/gahhg.h"

int main(int argc, char** argv) {
  std::string poly = "NA0.txt", tp = "NA_query.csv", report = "report.txt", result = "result.csv";
  int maxdepth = -1;
  for (int i = 1; i < argc; i++) {
    auto val = [&](std::string* dst) {
      if (i + 1 < argc) *dst = argv[++i];
    };
    if (!strcmp(argv[i], "/polyfile")) val(&poly);
    else if (!strcmp(argv[i], "/reportfile")) val(&report);
    else if (!strcmp(argv[i], "/resultfile")) val(&result);
    else if (!strcmp(argv[i], "/tpfile")) val(&tp);
    else if (!strcmp(argv[i], "/maxdepth") && i + 1 < argc) maxdepth = atoi(argv[++i]);
  }
  gahhg_params prm;
  gahhg_params_default(&prm);
  prm.verbose = 1;
  if (maxdepth > 0) {
    prm.max_depth = maxdepth;
    prm.auto_depth = 0;
  }
  gahhg_ctx* ctx = nullptr;
  if (gahhg_create(&ctx, &prm) != GAHHG_OK) {
    fprintf(stderr, "Error. No usable GPU.\n");
    return -1;
  }
  // polygon and query text are parsed on the GPU (values equal the host
  // loaders' on every line; the query points stay in device memory)
  printf("Load spherical polygon from %s.\n", poly.c_str());
  if (gahhg_load_polygon_txt_device(ctx, poly.c_str(), nullptr) != GAHHG_OK) {
    printf("Error. %s\n", gahhg_last_error(ctx));
    gahhg_destroy(ctx);
    return -1;
  }
  printf("Load test points from a CSV file: %s.\n", tp.c_str());
  const double* d_xyz = nullptr;
  int64_t n = 0;
  if (gahhg_load_query_csv_device(ctx, tp.c_str(), &d_xyz, &n, nullptr) != GAHHG_OK) {
    printf("Error. Can not read %s\n", tp.c_str());
    gahhg_destroy(ctx);
    return -1;
  }
  std::vector<int32_t> labels((size_t)n);
  void* d_labels = nullptr;
  if (gahhg_build(ctx) != GAHHG_OK || gahhg_device_alloc(ctx, std::max<int64_t>(n, 1) * 4, &d_labels) != GAHHG_OK ||
      gahhg_query_xyz_device(ctx, d_xyz, n, (int32_t*)d_labels) != GAHHG_OK ||
      gahhg_memcpy_d2h(ctx, labels.data(), d_labels, n * 4) != GAHHG_OK) {
    printf("Error. %s\n", gahhg_last_error(ctx));
    gahhg_destroy(ctx);
    return -1;
  }
  gahhg_device_free(ctx, d_labels);
  auto base = [](const std::string& s) {
    size_t k = s.find_last_of("/\\");
    return k == std::string::npos ? s : s.substr(k + 1);
  };
  gahhg_write_report(ctx, report.c_str(), base(poly).c_str(), base(tp).c_str());
  gahhg_write_result_csv(result.c_str(), labels.data(), n);
  gahhg_destroy(ctx);
  return 0;
}
