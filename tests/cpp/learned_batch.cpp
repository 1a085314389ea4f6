// Batched learned-init registration through rtr::LearnedRegistration
// (include/rtr_registration.hpp): the reference's load -> coarse alignment ->
// ICP -> transform sequence (RealTimeRobot.cpp:34-35,45-105;
// function.h:111-123) for many (model, scan) pairs at once, written against
// the drop-in header.  Prints one line per pair: "pair <i> rejected <0|1>
// iterations <n> T" followed by the 16 column-major doubles of the final
// transform (%.17g).
// Usage: learned_batch <weights.f32> <src0.pcd> <tgt0.pcd> [<src1.pcd> <tgt1.pcd> ...]
#include <cstdio>
#include <iostream>
#include <memory>
#include <vector>

#include "rtr_registration.hpp"

int main(int argc, char** argv) {
  if (argc < 4 || (argc - 2) % 2 != 0) {
    std::cerr << "usage: learned_batch weights.f32 src0.pcd tgt0.pcd [src1.pcd tgt1.pcd ...]\n";
    return 2;
  }
  using Cloud = rtr::PointCloud<rtr::PointXYZ>;
  std::vector<std::shared_ptr<const Cloud>> models, scans;
  for (int a = 2; a + 1 < argc; a += 2) {
    auto m = std::make_shared<Cloud>(), s = std::make_shared<Cloud>();
    if (rtr::io::loadPCDFile(argv[a], *m) != 0 || rtr::io::loadPCDFile(argv[a + 1], *s) != 0) {
      std::cerr << "load failed: " << rtr_last_error() << "\n";
      return 1;
    }
    models.push_back(m);
    scans.push_back(s);
  }
  try {
    rtr::LearnedRegistration<rtr::PointXYZ> reg((int)models.size());
    reg.loadWeights(argv[1]);
    reg.setInputClouds(models, scans);
    reg.align();
    const rtr_learned_stats st = reg.getStats();
    for (size_t i = 0; i < models.size(); ++i) {
      const rtr::Matrix4d T = reg.getFinalTransformation(i);
      std::printf("pair %zu rejected %d iterations %d T", i, reg.getRejected()[i], reg.getResults()[i].iterations);
      for (int k = 0; k < 16; ++k) std::printf(" %.17g", T.m[k]);
      std::printf("\n");
    }
    std::printf("pairs %d fallbacks %d total_ms %.2f\n", st.n_pairs, st.n_fallback, st.ms_total);
    return 0;
  } catch (const rtr::Error& e) {
    std::cerr << e.what() << "\n";
    return 3;
  }
}
