// Host-side AddressSanitizer / UndefinedBehaviorSanitizer driver.
//
// Built (tests/cpp/Makefile, target `sanitize`) against the library compiled
// with host-only ASan + UBSan (manybody_amd/Makefile, target `sanitize`): the
// device code is not instrumented, every host path is.  It drives the C ABI
// through the paths with host-side memory and threading risk:
//  * host only (always): bases, index lookups, the reference-order DFS tables,
//    error paths with null and out-of-range arguments, the tridiagonal solver,
//    KPM kernels, the rescaling and conductivity helpers, the sparse-matrix
//    container and its cereal file round trip, request footprints;
//  * with a GPU: operators and composites created and destroyed, concurrent
//    H.v on one operator from four host threads and streams (compared with the
//    serial result bit for bit), the batched engine (graph capture and
//    replay), the serving loop with the Lanczos lane fed from four submitting
//    threads while four others wait (the arena carving and the worker
//    threads), a request larger than the arena, and the no-allocation Lanczos.
// Prints "HOST PASSED" after the host part and "ALL PASSED" after the device
// part; any sanitizer report aborts the run with a nonzero status.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <unistd.h>
#include <vector>

#include "manybody_hip.h"

static std::atomic<int> failures{0};
#define CHECK(cond)                                                          \
  do {                                                                       \
    if (!(cond)) {                                                           \
      std::printf("FAIL %s:%d: %s (%s)\n", __FILE__, __LINE__, #cond, mb_last_error()); \
      ++failures;                                                            \
    }                                                                        \
  } while (0)

static void host_part() {
  // bases and lookups
  mb_basis *b = nullptr;
  CHECK(mb_basis_create(10, 9, 1, &b) == MB_OK);  // 5 up, 4 down
  CHECK(mb_basis_size(b) == 252 * 210);
  for (int64_t i = 0; i < mb_basis_size(b); i += 997) {
    uint64_t u = 0, d = 0;
    int64_t j = -1;
    CHECK(mb_basis_state(b, i, &u, &d) == MB_OK);
    CHECK(mb_basis_index_of(b, u, d, &j) == MB_OK && j == i);
  }
  int64_t j = 0;
  CHECK(mb_basis_index_of(b, 0x3u, 0x1u, &j) == MB_ERR_RANGE);
  std::vector<int64_t> pidx(100);
  CHECK(mb_basis_product_index(b, 5, 100, pidx.data()) == MB_OK);
  CHECK(mb_basis_product_index(b, mb_basis_size(b) - 1, 2, pidx.data()) != MB_OK);
  mb_basis *fixed = nullptr, *mom = nullptr;
  CHECK(mb_fixed_n_basis_create(8, 5, &fixed) == MB_OK);
  const int sz[1] = {8}, K[1] = {2};
  CHECK(mb_momentum_basis_create(8, 6, 0, 1, sz, K, &mom) == MB_OK);
  CHECK(mb_basis_size(mom) > 0);
  mb_basis_destroy(fixed);
  mb_basis_destroy(mom);
  mb_basis *too_big = nullptr;
  CHECK(mb_basis_create(40, 4, 0, &too_big) == MB_ERR_INVALID);  // > 32 orbitals refused
  mb_basis_destroy(b);
  // errors on null handles and arguments
  CHECK(mb_basis_state(nullptr, 0, nullptr, nullptr) != MB_OK);
  CHECK(mb_hv_c128(nullptr, nullptr, nullptr, nullptr) != MB_OK);
  CHECK(mb_lanczos_pass_one(nullptr, nullptr, 5, nullptr, nullptr, nullptr, nullptr, nullptr) != MB_OK);
  CHECK(mb_lanczos_pass_one_noalloc(nullptr, nullptr, 5, nullptr, nullptr, nullptr, nullptr, nullptr) != MB_OK);
  CHECK(std::strlen(mb_last_error()) > 0);
  // tridiagonal solver (Toeplitz closed form), kernels and helpers
  const int m = 40;
  std::vector<double> a(m, 1.0), be(m, 0.5), w(m), V(m * m);
  CHECK(mb_tridiag_eig(m, a.data(), be.data(), w.data(), V.data()) == MB_OK);
  for (int k = 1; k <= m; ++k) {
    const double ex = 1.0 + 2.0 * 0.5 * std::cos(k * M_PI / (m + 1));
    double best = 1e9;
    for (double x : w) best = std::fmin(best, std::fabs(x - ex));
    CHECK(best < 1e-12);
  }
  std::vector<double> g(257);
  CHECK(mb_jackson_kernel(256, g.data()) == MB_OK && std::fabs(g[0] - 1.0) < 1e-15);
  std::vector<double> mu(65), lk(65);  // M + 1 and N + 1 values
  CHECK(mb_projector_moments(64, 0.2, mu.data()) == MB_OK);
  CHECK(mb_lorentz_kernel(64, 4.0, lk.data()) == MB_OK);
  double ra = 0, rb = 0;
  CHECK(mb_rescaling_from_bounds(-3.0, 5.0, 0.01, &ra, &rb) == MB_OK && std::fabs(rb - 1.0) < 1e-15);
  std::vector<double> corr(2 * 50);
  for (int i = 0; i < 50; ++i) corr[2 * i] = std::exp(-0.1 * i);
  std::vector<double> freq(512), sig(2 * 512);  // sigma: (re, im) per frequency
  int nf = 0;
  double rule = 0;
  CHECK(mb_optical_conductivity(corr.data(), 50, 0.1, 2.0, 4.0, freq.data(), sig.data(), 512, &nf,
                                &rule) == MB_OK && nf > 0);
  // sparse container and its cereal round trip
  const uint64_t rows[3] = {0, 2, 1}, cols[3] = {0, 0, 2};
  const double vals[6] = {1.0, 0.5, -2.0, 0.0, 3.0, -1.0};
  mb_sparse *sp = nullptr, *sp2 = nullptr;
  CHECK(mb_sparse_create(3, 3, 3, rows, cols, vals, &sp) == MB_OK);
  const std::string path = std::string("/tmp/mb_sanitize_") + std::to_string(getpid()) + ".bin";
  CHECK(mb_sparse_save_cereal(sp, path.c_str()) == MB_OK);
  CHECK(mb_sparse_load_cereal(path.c_str(), &sp2) == MB_OK);
  uint64_t nr = 0, nc = 0;
  int64_t nnz = 0;
  CHECK(mb_sparse_shape(sp2, &nr, &nc, &nnz) == MB_OK && nr == 3 && nc == 3 && nnz == 3);
  std::vector<uint64_t> r2(3), c2(3);
  std::vector<double> v2(6);
  CHECK(mb_sparse_entries(sp2, r2.data(), c2.data(), v2.data()) == MB_OK);
  CHECK(r2[1] == 2 && c2[2] == 2 && v2[5] == -1.0);
  mb_sparse_destroy(sp2);
  std::remove(path.c_str());
  mb_sparse *missing = nullptr;
  CHECK(mb_sparse_load_cereal(path.c_str(), &missing) != MB_OK);
  mb_sparse_destroy(sp);
  // request footprints
  mb_request r{};
  r.kind = MB_REQ_KPM;
  r.ndim = 1;
  r.sizes[0] = 12;
  r.n_up = 6;
  r.n_dn = 6;
  r.t = 1.0;
  r.steps = 40;
  r.moments = 64;
  r.nvec = 4;
  r.arrival = -1;
  CHECK(mb_request_footprint(&r) > 0);
  r.n_up = 13;  // more particles than orbitals
  CHECK(mb_request_footprint(&r) == -1);
  std::printf("HOST PASSED\n");
}

static mb_hubbard *ring(mb_basis **keep, int L, int nu, int nd, double U) {
  CHECK(mb_basis_create(L, nu + nd, nu - nd, keep) == MB_OK);
  mb_hubbard *h = nullptr;
  const int s[1] = {L};
  CHECK(mb_hubbard_create(*keep, 1, s, 1.0, U, &h) == MB_OK);
  return h;
}

static void device_part() {
  // concurrent H.v on one operator, four threads and streams
  mb_basis *b = nullptr;
  mb_hubbard *h = ring(&b, 12, 6, 5, 4.0);
  const int64_t D = mb_hubbard_dimension(h);
  double *x = nullptr, *ys = nullptr;
  CHECK(hipMalloc(&x, 16 * D) == hipSuccess && hipMalloc(&ys, 5 * 16 * D) == hipSuccess);
  CHECK(mb_fill_gaussian_c128(h, 7, x, nullptr) == MB_OK);
  CHECK(mb_hv_c128(h, x, ys, nullptr) == MB_OK);
  CHECK(hipDeviceSynchronize() == hipSuccess);
  std::vector<std::thread> th;
  for (int t = 0; t < 4; ++t)
    th.emplace_back([&, t] {
      hipStream_t s;
      if (hipStreamCreate(&s) != hipSuccess) return;
      for (int rep = 0; rep < 20; ++rep) CHECK(mb_hv_c128(h, x, ys + (t + 1) * 2 * D, s) == MB_OK);
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    });
  for (auto &t : th) t.join();
  std::vector<double> h0(2 * D), h1(2 * D);
  CHECK(hipMemcpy(h0.data(), ys, 16 * D, hipMemcpyDeviceToHost) == hipSuccess);
  for (int t = 1; t <= 4; ++t) {
    CHECK(hipMemcpy(h1.data(), ys + t * 2 * D, 16 * D, hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(std::memcmp(h0.data(), h1.data(), 16 * D) == 0);
  }
  // composites and the no-allocation Lanczos
  mb_hubbard *neg = nullptr, *sum = nullptr;
  CHECK(mb_operator_negated(h, &neg) == MB_OK);
  CHECK(mb_operator_sum(h, neg, &sum) == MB_OK);
  CHECK(mb_hv_c128(sum, x, ys, nullptr) == MB_OK);
  std::vector<double> al(30), bt(30), al2(30), bt2(30);
  int st = 0, st2 = 0;
  double bn = 0, bn2 = 0;
  CHECK(mb_lanczos_pass_one(h, x, 30, al.data(), bt.data(), &st, &bn, nullptr) == MB_OK);
  CHECK(mb_lanczos_pass_one_noalloc(h, x, 30, al2.data(), bt2.data(), &st2, &bn2, nullptr) == MB_OK);
  CHECK(st == st2 && std::memcmp(al.data(), al2.data(), 8 * 30) == 0);
  mb_hubbard_destroy(sum);
  mb_hubbard_destroy(neg);
  CHECK(hipFree(x) == hipSuccess && hipFree(ys) == hipSuccess);
  // batched engine, twice (graph captured, replayed, destroyed)
  for (int round = 0; round < 2; ++round) {
    std::vector<mb_basis *> bs(24);
    std::vector<mb_hubbard *> hs(24);
    for (int i = 0; i < 24; ++i) hs[i] = ring(&bs[i], 8 + i % 5, 1 + i % 4, 1 + (i / 3) % 4, 0.5 * i);
    mb_batch *bat = nullptr;
    CHECK(mb_batch_create(24, hs.data(), &bat) == MB_OK);
    std::vector<uint64_t> seeds(24);
    for (int i = 0; i < 24; ++i) seeds[i] = 100 + i;
    CHECK(mb_batch_fill_gaussian(bat, seeds.data(), nullptr) == MB_OK);
    std::vector<double> e0(24), e1(24);
    CHECK(mb_batch_ground_energies(bat, 60, 1, e0.data(), nullptr) == MB_OK);
    CHECK(mb_batch_ground_energies(bat, 60, 1, e1.data(), nullptr) == MB_OK);
    CHECK(std::memcmp(e0.data(), e1.data(), 8 * 24) == 0);
    mb_batch_destroy(bat);
    for (int i = 0; i < 24; ++i) {
      mb_hubbard_destroy(hs[i]);
      mb_basis_destroy(bs[i]);
    }
  }
  // serving loop: four submitting threads, four waiting threads, the lane
  mb_server *srv = nullptr;
  // 512 MB arena, 128 MB of it for the lane: the L = 14, 7 + 7 request below does not fit
  CHECK(mb_server_create(3, int64_t(512) << 20, int64_t(128) << 20, &srv) == MB_OK);
  std::vector<int64_t> ids(48, -1);
  std::vector<mb_result> res(48);
  std::vector<std::thread> sub, wt;
  std::atomic<int> submitted{0};
  for (int t = 0; t < 4; ++t)
    sub.emplace_back([&, t] {
      for (int i = t; i < 48; i += 4) {
        mb_request q{};
        q.kind = i % 5 == 3 ? MB_REQ_KPM : (i % 7 == 5 ? MB_REQ_EVOLVE : MB_REQ_GROUND);
        q.ndim = 1;
        q.sizes[0] = 8 + i % 4;
        q.n_up = 1 + i % 4;
        q.n_dn = 1 + (i / 4) % 4;
        q.t = 1.0;
        q.u = 0.25 * i;
        q.steps = q.kind == MB_REQ_EVOLVE ? 20 : 60;
        q.moments = 32;
        q.nvec = 2;
        q.t_final = 0.3;
        q.dt = 0.05;
        q.seed = 500 + i;
        q.arrival = -1.0;
        int64_t id = -1;
        CHECK(mb_server_submit(srv, &q, &id) == MB_OK);
        ids[i] = id;
        submitted++;
      }
    });
  for (auto &t : sub) t.join();
  for (int t = 0; t < 4; ++t)
    wt.emplace_back([&, t] {
      std::vector<double> mom(64);
      for (int i = t; i < 48; i += 4) CHECK(mb_server_wait(srv, ids[i], &res[i], mom.data(), 64) == MB_OK);
    });
  for (auto &t : wt) t.join();
  for (int i = 0; i < 48; ++i) CHECK(res[i].status == MB_OK);
  mb_request big{};
  big.kind = MB_REQ_GROUND;
  big.ndim = 1;
  big.sizes[0] = 14;
  big.n_up = 7;
  big.n_dn = 7;
  big.t = 1.0;
  big.steps = 40;
  big.arrival = -1.0;
  int64_t bid = -1;
  CHECK(mb_server_submit(srv, &big, &bid) == MB_OK);
  mb_result br{};
  CHECK(mb_server_wait(srv, bid, &br, nullptr, 0) == MB_OK && br.status == MB_ERR_NOMEM);
  CHECK(mb_server_drain(srv) == MB_OK);
  mb_server_destroy(srv);
  mb_hubbard_destroy(h);
  mb_basis_destroy(b);
  std::printf("ALL PASSED\n");
}

int main() {
  std::setvbuf(stdout, nullptr, _IONBF, 0);
  host_part();
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    std::printf("no GPU: device part skipped\n");
  } else {
    device_part();
  }
  if (failures) std::printf("%d FAILURES\n", failures.load());
  // leave without running static destructors: the HIP runtime's own teardown
  // frees memory through the sanitizer's device allocator, which rejects it
  std::fflush(stdout);
  _exit(failures ? 1 : 0);
}
