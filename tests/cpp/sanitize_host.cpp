// The library's host paths under AddressSanitizer and UndefinedBehaviorSanitizer
// (built by tests/cpp/Makefile.sanitize, run by tests/test_host_sanitizers.py):
// the PCD v0.7 reader and writer (pcd.cpp) on the golden files and on
// malformed headers, the self-check hooks (selfcheck.hip: the device
// geometry code of geom.hpp run on the host) and the CPU oracle (oracle.c),
// on random and tie-heavy clouds.  A sanitizer finding aborts the run; a
// wrong answer prints what differed and exits with status 1.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "oracle.h"
#include "rtr.h"

static int g_fail = 0;

#define EXPECT(cond, ...)                          \
  do {                                             \
    if (!(cond)) {                                 \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);                \
      fprintf(stderr, "\n");                       \
      ++g_fail;                                    \
    }                                              \
  } while (0)

static std::vector<float> random_cloud(std::mt19937& rng, size_t n, float lo, float hi, float grid) {
  std::uniform_real_distribution<float> u(lo, hi);
  std::vector<float> p(3 * n);
  for (auto& v : p) {
    v = u(rng);
    if (grid > 0) v = std::round(v / grid) * grid;  // grid-aligned: exact distance ties
  }
  return p;
}

static void check_pcd(const std::vector<std::string>& files, const std::string& tmp) {
  for (const auto& f : files) {
    size_t n = 0;
    int hn = 0;
    int rc = rtr_pcd_info(f.c_str(), &n, &hn);
    EXPECT(rc == RTR_OK, "pcd_info %s: %d %s", f.c_str(), rc, rtr_last_error());
    if (rc) continue;
    std::vector<float> xyz(3 * (n ? n : 1)), nrm(3 * (n ? n : 1));
    size_t got = 0;
    rc = rtr_pcd_read(f.c_str(), xyz.data(), hn ? nrm.data() : nullptr, n, &got);
    EXPECT(rc == RTR_OK && got == n, "pcd_read %s: %d", f.c_str(), rc);
    // a buffer one point short is a capacity error, not an overflow
    if (n > 0) {
      std::vector<float> small(3 * (n - 1) + 1);
      size_t g2 = 0;
      rc = rtr_pcd_read(f.c_str(), small.data(), nullptr, n - 1, &g2);
      EXPECT(rc == RTR_ERR_CAPACITY, "short buffer for %s: %d", f.c_str(), rc);
    }
    // ascii round trip through the writer (8 significant digits)
    const std::string out = tmp + "/rt.pcd";
    rc = rtr_pcd_write_ascii(out.c_str(), xyz.data(), n, 12);
    EXPECT(rc == RTR_OK, "write %s", out.c_str());
    std::vector<float> back(3 * (n ? n : 1));
    size_t g3 = 0;
    rc = rtr_pcd_read(out.c_str(), back.data(), nullptr, n, &g3);
    EXPECT(rc == RTR_OK && g3 == n, "re-read %s", out.c_str());
    for (size_t i = 0; i < 3 * n && rc == RTR_OK; ++i) {
      const float a = xyz[i], b = back[i];
      EXPECT((std::isnan(a) && std::isnan(b)) || std::fabs(a - b) <= 1e-7f * std::fmax(1.f, std::fabs(a)),
             "round trip %s value %zu: %g vs %g", f.c_str(), i, a, b);
      if (!((std::isnan(a) && std::isnan(b)) || std::fabs(a - b) <= 1e-7f * std::fmax(1.f, std::fabs(a)))) break;
    }
  }
  // malformed headers: a format error, never a read past the payload
  const char* heads[] = {
      "FIELDS x y z\nSIZE 4 4 4\nTYPE F F F\nCOUNT 1 1\nWIDTH 2\nHEIGHT 1\nPOINTS 2\nDATA binary\n",
      "FIELDS x y z\nSIZE 2 4 4\nTYPE F F F\nCOUNT 1 1 1\nWIDTH 2\nHEIGHT 1\nPOINTS 2\nDATA binary\n",
      "FIELDS x y z\nSIZE 4 4 4\nTYPE F F F\nCOUNT 1 0 1\nWIDTH 2\nHEIGHT 1\nPOINTS 2\nDATA binary\n",
      "FIELDS x y z\nSIZE 4 4 -8\nTYPE F F F\nCOUNT 1 1 1\nWIDTH 2\nHEIGHT 1\nPOINTS 2\nDATA binary\n",
      "FIELDS x y z\nSIZE 8 8 8\nTYPE F F F\nCOUNT 1 1 1\nWIDTH 2\nHEIGHT 1\nPOINTS 4611686018427387905\nDATA binary\n",
      "FIELDS x y z\nSIZE 4 4 4\nTYPE F F F\nCOUNT 1 1 1\nWIDTH 4294967296\nHEIGHT 4294967297\nDATA binary\n",
      "FIELDS x y z\nSIZE 4 4 4\nTYPE F F F\nCOUNT 1 1 1\nWIDTH 3\nHEIGHT 1\nPOINTS 3\nDATA binary\n",
  };
  const std::string bad = tmp + "/bad.pcd";
  for (const char* h : heads) {
    FILE* fp = fopen(bad.c_str(), "wb");
    fputs(h, fp);
    const unsigned char payload[24] = {0};
    fwrite(payload, 1, sizeof(payload), fp);  // 2 points of 12 bytes: too few for 3
    fclose(fp);
    size_t n = 0;
    int hn = 0;
    const int rc = rtr_pcd_info(bad.c_str(), &n, &hn);
    EXPECT(rc == RTR_ERR_FORMAT, "malformed header accepted (%d): %s", rc, h);
  }
}

static void check_nn() {
  std::mt19937 rng(7);
  for (int trial = 0; trial < 6; ++trial) {
    const float grid = trial % 2 ? 0.01f : 0.f;
    auto t = random_cloud(rng, 3000, 0.f, 0.6f, grid);
    auto q = random_cloud(rng, 1500, -0.1f, 0.7f, grid);
    if (trial == 3) t[3 * 17] = NAN;
    const double rmax[3] = {0.02, 0.05, std::sqrt(1e300)};
    for (double r : rmax) {
      const float r2f = orc_r2_to_float(r * r);
      std::vector<int32_t> ib(1500), is(1500), ik(1500);
      std::vector<float> db(1500), ds(1500), dk(1500);
      orc_nn_brute(q.data(), 1500, t.data(), 3000, r2f, ib.data(), db.data());
      const float cells[3] = {0.011f, 0.03f, 0.2f};
      for (float cell : cells) {
        int rc = rtr_selfcheck_nn(t.data(), 3000, q.data(), 1500, cell, r, is.data(), ds.data());
        EXPECT(rc == RTR_OK, "selfcheck_nn rc %d", rc);
        EXPECT(memcmp(ib.data(), is.data(), 1500 * 4) == 0 && memcmp(db.data(), ds.data(), 1500 * 4) == 0,
               "grid NN differs from brute force (trial %d, r %g, cell %g)", trial, r, cell);
      }
      orc_kdtree* kd = orc_kd_build(t.data(), 3000, 15);
      orc_kd_nn(kd, q.data(), 1500, r2f, ORC_KD_TIE_LOWEST | ORC_KD_BOUNDED, ik.data(), dk.data());
      EXPECT(memcmp(ib.data(), ik.data(), 1500 * 4) == 0, "bounded kd NN differs (trial %d, r %g)", trial, r);
      std::vector<int32_t> kidx(1500 * 8);
      std::vector<float> kd2(1500 * 8);
      orc_kd_knn(kd, q.data(), 1500, 8, kidx.data(), kd2.data());
      for (int i = 0; i < 1500; ++i)
        if (ib[i] >= 0 && db[i] < INFINITY) EXPECT(kd2[8 * i] == db[i], "knn first != nn at %d", i);
      orc_kd_free(kd);
    }
  }
}

static void check_solves() {
  std::mt19937 rng(11);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int trial = 0; trial < 20; ++trial) {
    auto a = random_cloud(rng, 500, -1.f, 1.f, 0.f);
    // b = R a + t exactly in double, then float
    double ax[3] = {nd(rng), nd(rng), nd(rng)};
    const double nrm = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    const double th = 0.3 * nd(rng);
    const double k[3] = {ax[0] / nrm, ax[1] / nrm, ax[2] / nrm};
    const double c = std::cos(th), s = std::sin(th), v = 1 - c;
    const double R[9] = {c + k[0] * k[0] * v,        k[0] * k[1] * v - k[2] * s, k[0] * k[2] * v + k[1] * s,
                         k[1] * k[0] * v + k[2] * s, c + k[1] * k[1] * v,        k[1] * k[2] * v - k[0] * s,
                         k[2] * k[0] * v - k[1] * s, k[2] * k[1] * v + k[0] * s, c + k[2] * k[2] * v};
    const double t[3] = {0.1 * nd(rng), 0.1 * nd(rng), 0.1 * nd(rng)};
    std::vector<float> b(a.size());
    std::vector<double> ad(a.size()), bd(a.size());
    for (size_t i = 0; i < a.size() / 3; ++i)
      for (int r = 0; r < 3; ++r) {
        b[3 * i + r] = (float)(R[3 * r] * a[3 * i] + R[3 * r + 1] * a[3 * i + 1] + R[3 * r + 2] * a[3 * i + 2] + t[r]);
        ad[3 * i + r] = a[3 * i + r];
        bd[3 * i + r] = b[3 * i + r];
      }
    double Ts[16], To[16];
    const float anchor[3] = {0.25f, -0.5f, 1.0f};
    int rc = rtr_selfcheck_solve(RTR_MODE_POINT_TO_POINT, a.data(), b.data(), nullptr, a.size() / 3, anchor, Ts);
    EXPECT(rc == RTR_OK, "selfcheck_solve rc %d", rc);
    orc_umeyama(ad.data(), bd.data(), (int64_t)(a.size() / 3), To);
    double err = 0;
    for (int i = 0; i < 16; ++i) err = std::fmax(err, std::fabs(Ts[i] - To[i]));
    EXPECT(err < 1e-9, "rigid solve differs from the oracle by %g (trial %d)", err, trial);
  }
  // Anderson mixing on an affine contraction: converges to (I - A)^-1 c
  double A[36] = {0}, cv[6] = {1, -2, 0.5, 0.25, 3, -1}, x[6];
  for (int i = 0; i < 6; ++i) A[7 * i] = 0.9 - 0.1 * i;
  int iters = 0;
  int rc = rtr_selfcheck_anderson(6, A, cv, 1e-12, 200, x, &iters);
  EXPECT(rc == RTR_OK, "anderson rc %d", rc);
  for (int i = 0; i < 6; ++i) EXPECT(std::fabs(x[i] - cv[i] / (1.0 - A[7 * i])) < 1e-8, "anderson x[%d]", i);
  const double v[3] = {0.1, -0.2, 0.3};
  double Rm[9], vb[3];
  rc = rtr_selfcheck_so3(v, Rm, vb);
  EXPECT(rc == RTR_OK && std::fabs(vb[0] - v[0]) < 1e-12 && std::fabs(vb[2] - v[2]) < 1e-12, "so3 round trip");
  const double vbad[3] = {2.0, 0, 0};
  EXPECT(rtr_selfcheck_so3(vbad, Rm, vb) == RTR_ERR_INVALID_ARG, "so3 outside the chart accepted");
}

static void check_oracle_misc() {
  std::mt19937 rng(3);
  auto t = random_cloud(rng, 4000, 0.f, 0.5f, 0.f);
  auto s = random_cloud(rng, 2000, 0.f, 0.5f, 0.f);
  orc_icp_params p;
  memset(&p, 0, sizeof(p));
  p.max_iter = 10;
  p.min_corr = 3;
  p.r_max = 0.05;
  p.eps_T = 1e-8;
  p.eps_fit = -1e300;
  orc_icp_result r1, r2;
  orc_icp(s.data(), 2000, t.data(), nullptr, 4000, &p, nullptr, &r1, nullptr);
  orc_pcl_icp(s.data(), 2000, t.data(), 4000, &p, nullptr, &r2);
  EXPECT(r1.iterations >= 1 && r2.iterations >= 1, "icp ran no iteration");
  const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  const double f = orc_fitness(s.data(), 2000, t.data(), 4000, I, 1e300);
  EXPECT(f >= 0.0, "fitness %g", f);
  std::vector<float> nrm(3 * 4000);
  const float vp[3] = {0, 0, 0};
  orc_normals(t.data(), 4000, 20, vp, nrm.data(), nullptr);
  std::vector<float> down(3 * 4000);
  const int64_t nd = orc_voxel_downsample(t.data(), 4000, 0.05f, down.data(), 4000, nullptr);
  EXPECT(nd > 0 && nd <= 4000, "downsample count %lld", (long long)nd);
  std::vector<int32_t> sel(256);
  double cen[3];
  orc_fps(t.data(), 4000, 8192, 256, sel.data(), cen);
  for (int i = 0; i < 256; ++i) EXPECT(sel[i] >= 0 && sel[i] < 4000, "fps pick %d", sel[i]);
  const int dims[3] = {12, 9, 7};
  const int32_t occ[12] = {0, 0, 0, 5, 4, 3, 11, 8, 6, 40, 0, 0};  // the last lies outside the grid
  std::vector<float> tdf(12 * 9 * 7);
  orc_tdf(occ, 4, dims, 900.f, tdf.data());
  EXPECT(tdf[0] == 0.f && tdf[5 + 4 * 12 + 3 * 108] == 0.f && tdf[1] == 1.f, "tdf at occupied voxels");
}

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s TMPDIR [PCD files...]\n", argv[0]);
    return 2;
  }
  std::vector<std::string> files(argv + 2, argv + argc);
  check_pcd(files, argv[1]);
  check_nn();
  check_solves();
  check_oracle_misc();
  if (g_fail) {
    fprintf(stderr, "%d check(s) failed\n", g_fail);
    return 1;
  }
  printf("sanitized host paths: %zu PCD files, grid/kd NN, solves, Anderson, oracle: OK\n", files.size());
  return 0;
}
