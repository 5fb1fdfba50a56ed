// C ABI of libopentbl_amd: box setup, state transfer, the RK3 substep (eager
// or as captured hipGraphs), operator entry points and kernel timing.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <mutex>
#include <string>

#include "otbl_internal.h"
#include "otbl_ycol.h"
#include "otbl_zfft.h"

namespace otbl {

static thread_local std::string g_err;

// Scratch of otbl_soltriy_noalloc: the caller's (3, 1:ny1) tridiagonal of
// one call, at most kNoAllocRows rows.  A file-scope device array: each GPU
// the library runs on holds its own copy until the process ends, and every
// call on that GPU shares it.  Host code reaches it only through the
// per-device address hipGetSymbolAddress returns (the array's name is a
// host-side stand-in), looked up once per device at otbl_create.
constexpr int kNoAllocRows = 8192;
static __device__ __align__(256) double scratch[3 * (kNoAllocRows + 1)];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch[kMaxDevices] = {};

// the address of this device's scratch, set up on first use with the device
// current (called by otbl_create)
static double *scratch_of(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw std::invalid_argument("device index above 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    OTBL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("hipGetSymbolAddress returned a null scratch address");
    g_scratch[dev] = static_cast<double *>(p);
  }
  return g_scratch[dev];
}

void fail_hip(hipError_t e, const char *what) {
  throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e) + " in " + what);
}

void kbegin(Box &b, const char *name) {
  if (!b.kprof) return;
  hipEvent_t e0, e1;
  OTBL_HIP(hipEventCreate(&e0));
  OTBL_HIP(hipEventCreate(&e1));
  OTBL_HIP(hipEventRecord(e0, b.s));
  b.pending.push_back({name, {e0, e1}});
}

void kend(Box &b) {
  OTBL_HIP(hipGetLastError());
  if (!b.kprof) return;
  OTBL_HIP(hipEventRecord(b.pending.back().second.second, b.s));
}

static void drain_kprof(Box &b) {
  if (b.pending.empty()) return;
  OTBL_HIP(hipStreamSynchronize(b.s));
  for (auto &p : b.pending) {
    float ms = 0.f;
    OTBL_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
    auto &st = b.kstat[p.first];
    st.n += 1;
    st.ms += ms;
    OTBL_HIP(hipEventDestroy(p.second.first));
    OTBL_HIP(hipEventDestroy(p.second.second));
  }
  b.pending.clear();
}

static int small_factors(int n, int *fac) {
  int nf = 0;
  const int pr[6] = {8, 4, 2, 3, 5, 7};
  for (int q = 0; q < 6; ++q)
    while (n % pr[q] == 0) {
      fac[nf++] = pr[q];
      n /= pr[q];
    }
  return n == 1 ? nf : -1;
}

template <class T> static void dalloc(T *&p, size_t n) {
  OTBL_HIP(hipMalloc((void **)&p, n * sizeof(T) + 64));
  OTBL_HIP(hipMemset(p, 0, n * sizeof(T) + 64));
  // the clear runs on the null stream, which the box's non-blocking stream
  // does not wait for
  OTBL_HIP(hipDeviceSynchronize());
}

// A field or work field of n elements of the box's element size (8 bytes
// in FP64, 4 in the mixed substep), held behind a double pointer.
static void falloc(const Box &b, double *&p, size_t n) { dalloc(p, (n * b.esz + 7) / 8); }

// Host <-> device copies ordered on the box's stream, complete on return.
// (The stream is non-blocking: a copy on the null stream is not ordered
// before kernels enqueued on it, and a pageable host-to-device hipMemcpy may
// return before its transfer lands.)
static void h2d(Box &b, void *dst, const void *src, size_t bytes) {
  OTBL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b.s));
  OTBL_HIP(hipStreamSynchronize(b.s));
}
static void d2h(Box &b, void *dst, const void *src, size_t bytes) {
  OTBL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, b.s));
  OTBL_HIP(hipStreamSynchronize(b.s));
}
static void put_ds(Box &b, size_t off, const void *src, size_t bytes) { h2d(b, (char *)b.ds + off, src, bytes); }
static void get_ds(Box &b, size_t off, void *dst, size_t bytes) { d2h(b, dst, (char *)b.ds + off, bytes); }

// Coefficient arrays, in upload order, with their device slot.  cols: the
// Fortran extent ('X' nx, 'Y' ny, 'U' ny+1, 'P' ny-1, 'M' nx-1, 'K' nz2+1,
// 'G' ny+2); first: the first Fortran index exported.
struct CoefSlot {
  const char *name;
  std::vector<double> HostCoef::*vec;
  const double *DevCoef::*dev;
  int rows;
  int first;
  char cols;
};

static const CoefSlot kSlots[] = {
    {"inxu", &HostCoef::inxu, &DevCoef::inxu, 3, 1, 'X'},       {"inxv", &HostCoef::inxv, &DevCoef::inxv, 3, 1, 'X'},
    {"inyu", &HostCoef::inyu, &DevCoef::inyu, 3, 1, 'Y'},       {"inyv", &HostCoef::inyv, &DevCoef::inyv, 3, 1, 'Y'},
    {"dcxu", &HostCoef::dcxu, &DevCoef::dcxu, 3, 1, 'X'},       {"dcxv", &HostCoef::dcxv, &DevCoef::dcxv, 3, 1, 'X'},
    {"dcyu", &HostCoef::dcyu, &DevCoef::dcyu, 3, 1, 'U'},       {"dcyv", &HostCoef::dcyv, &DevCoef::dcyv, 3, 1, 'Y'},
    {"vixu", &HostCoef::vixu, &DevCoef::vixu, 3, 1, 'X'},       {"vixv", &HostCoef::vixv, &DevCoef::vixv, 3, 1, 'X'},
    {"viyu", &HostCoef::viyu, &DevCoef::viyu, 3, 1, 'U'},       {"viyv", &HostCoef::viyv, &DevCoef::viyv, 3, 1, 'Y'},
    {"vyui", &HostCoef::vyui, &DevCoef::vyui, 3, 1, 'U'},       {"vyvi", &HostCoef::vyvi, &DevCoef::vyvi, 3, 1, 'Y'},
    {"cofiux", &HostCoef::cofiux, &DevCoef::cofiux, 2, 1, 'X'}, {"cofivx", &HostCoef::cofivx, &DevCoef::cofivx, 2, 1, 'X'},
    {"cofiuy", &HostCoef::cofiuy, &DevCoef::cofiuy, 2, 1, 'Y'}, {"cofivy", &HostCoef::cofivy, &DevCoef::cofivy, 2, 1, 'Y'},
    {"cofcxu", &HostCoef::cofcxu, &DevCoef::cofcxu, 2, 1, 'X'}, {"cofcxv", &HostCoef::cofcxv, &DevCoef::cofcxv, 2, 1, 'X'},
    {"cofcyu", &HostCoef::cofcyu, &DevCoef::cofcyu, 2, 1, 'Y'}, {"cofcyv", &HostCoef::cofcyv, &DevCoef::cofcyv, 2, 1, 'Y'},
    {"cofvxu", &HostCoef::cofvxu, &DevCoef::cofvxu, 3, 1, 'X'}, {"cofvxv", &HostCoef::cofvxv, &DevCoef::cofvxv, 3, 1, 'X'},
    {"cofvyu", &HostCoef::cofvyu, &DevCoef::cofvyu, 3, 1, 'Y'}, {"cofvyv", &HostCoef::cofvyv, &DevCoef::cofvyv, 3, 1, 'Y'},
    {"ayp", &HostCoef::ayp, &DevCoef::ayp, 3, 1, 'P'},
    {"y", &HostCoef::y, &DevCoef::y, 1, 0, 'G'},             {"dy", &HostCoef::dy, &DevCoef::dy, 1, 0, 'U'},
    {"idy", &HostCoef::idy, &DevCoef::idy, 1, 0, 'U'},       {"idyy", &HostCoef::idyy, &DevCoef::idyy, 1, 1, 'Y'},
    {"dymin", &HostCoef::dymin, &DevCoef::dymin, 1, 1, 'Y'}, {"hy", &HostCoef::hy, &DevCoef::hy, 1, 0, 'U'},
    {"kaz", &HostCoef::kaz, &DevCoef::kaz, 1, 0, 'K'},       {"kaz2", &HostCoef::kaz2, &DevCoef::kaz2, 1, 0, 'K'},
    {"kvis", &HostCoef::kvis, &DevCoef::kvis, 1, 0, 'K'},    {"kmod", &HostCoef::kmod, &DevCoef::kmod, 1, 1, 'M'},
    {"vmagic", &HostCoef::vmagic, &DevCoef::vmagic, 1, 1, 'X'},
};

static int slot_cols(const Box &b, char c) {
  switch (c) {
  case 'X': return b.nx;
  case 'Y': return b.ny;
  case 'U': return b.ny + 1;
  case 'P': return b.ny - 1;
  case 'M': return b.nx - 1;
  case 'K': return b.nz2 + 1;
  default: return b.ny + 2;
  }
}

static size_t fixed_len(const char *nm) {
  if (!strcmp(nm, "dcbx")) return 20;
  if (!strcmp(nm, "ccon")) return 8;
  return 16;
}

static void upload_coef(Box &b) {
  HostCoef &hc = b.hc;
  size_t total = 0;
  for (const auto &sl : kSlots) total += (hc.*(sl.vec)).size() + 8;
  total += 7 * 24;
  if (!b.coefbuf) dalloc(b.coefbuf, total);
  std::vector<double> host(total, 0.0);
  size_t off = 0;
  for (const auto &sl : kSlots) {
    const auto &v = hc.*(sl.vec);
    std::copy(v.begin(), v.end(), host.begin() + off);
    b.dc.*(sl.dev) = b.coefbuf + off;
    off += v.size() + 8;
  }
  struct { double *src; size_t n; const double *DevCoef::*dev; } fx[7] = {
      {hc.inbx, 16, &DevCoef::inbx}, {hc.inby, 16, &DevCoef::inby}, {hc.dcbx, 20, &DevCoef::dcbx},
      {hc.dcby, 16, &DevCoef::dcby}, {hc.ccon, 8, &DevCoef::ccon},  {hc.cofvbx, 16, &DevCoef::cofvbx},
      {hc.cofvby, 16, &DevCoef::cofvby}};
  for (auto &f : fx) {
    std::copy(f.src, f.src + f.n, host.begin() + off);
    b.dc.*(f.dev) = b.coefbuf + off;
    off += 24;
  }
  h2d(b, b.coefbuf, host.data(), total * sizeof(double));
}

static int good_length(int m) {
  for (;; ++m) {
    int x = m;
    for (int p : {2, 3, 5, 7})
      while (x % p == 0) x /= p;
    if (x == 1) return m;
  }
}

static FFTPlan make_plan(int n) {
  FFTPlan p{};
  p.n = n;
  p.nfac = small_factors(n, p.fac);
  if (p.nfac < 0 || p.nfac > 14) throw std::invalid_argument("transform length must factor into 2, 3, 5, 7");
  return p;
}

// Bluestein tables for the length-N real-to-real x transforms
// (ffts.f90:124-181): chirp exp(-i pi n^2/N), the transform of the
// conjugate chirp wrapped to length M (divided by M), and exp(-i pi k/2N).
static void build_dct_fft(Box &b) {
  const int N = b.nx - 1;
  // the register engine takes the shortest of its lengths that holds the
  // convolution (OTBL_DCT_POW2: the shortest power of two)
  int M2 = zdct_length(2 * N - 1);
  if (std::getenv("OTBL_DCT_POW2")) {
    M2 = 1;
    while (M2 < 2 * N - 1) M2 *= 2;
  }
  b.dct_zfast = zdct_supported(M2) && std::getenv("OTBL_Z_STOCKHAM") == nullptr;
  const int M = b.dct_zfast ? M2 : good_length(2 * N - 1);
  const double pi = b.hc.pi;
  b.dct_M = M;
  b.dct_plan = make_plan(M);
  std::vector<double> chirp(2 * (size_t)N), bm(2 * (size_t)M, 0.0), bf(2 * (size_t)M), tw4(2 * (size_t)N);
  for (int n = 0; n < N; ++n) {
    long q = ((long)n * n) % (2L * N);
    chirp[2 * n] = std::cos(pi * q / N);
    chirp[2 * n + 1] = -std::sin(pi * q / N);
  }
  for (int t = 0; t < N; ++t) {
    bm[2 * t] = chirp[2 * t];
    bm[2 * t + 1] = -chirp[2 * t + 1];
    if (t > 0) {
      bm[2 * (M - t)] = chirp[2 * t];
      bm[2 * (M - t) + 1] = -chirp[2 * t + 1];
    }
  }
  std::vector<double> ct(M), st(M);
  for (int t = 0; t < M; ++t) {
    ct[t] = std::cos(2.0 * pi * t / M);
    st[t] = std::sin(2.0 * pi * t / M);
  }
  for (int m = 0; m < M; ++m) {
    double sr = 0.0, si = 0.0;
    for (int t = 0; t < M; ++t) {
      double br = bm[2 * t], bi = bm[2 * t + 1];
      if (br == 0.0 && bi == 0.0) continue;
      int a = (int)(((long)m * t) % M);
      // exp(-2 pi i m t / M)
      sr += br * ct[a] + bi * st[a];
      si += bi * ct[a] - br * st[a];
    }
    bf[2 * m] = sr / M;
    bf[2 * m + 1] = si / M;
  }
  for (int k = 0; k < N; ++k) {
    tw4[2 * k] = std::cos(pi * k / (2.0 * N));
    tw4[2 * k + 1] = -std::sin(pi * k / (2.0 * N));
  }
  dalloc(b.dct_chirp, chirp.size());
  dalloc(b.dct_bf, bf.size());
  dalloc(b.dct_tw4, tw4.size());
  h2d(b, b.dct_chirp, chirp.data(), chirp.size() * sizeof(double));
  h2d(b, b.dct_bf, bf.data(), bf.size() * sizeof(double));
  h2d(b, b.dct_tw4, tw4.data(), tw4.size() * sizeof(double));
}

static void build_dct(Box &b) {
  const int N = b.nx - 1;
  std::vector<double> c2((size_t)N * N), c3((size_t)N * N);
  const double pi = b.hc.pi;
  // cos(pi t / (2N)) with t reduced mod 4N: the FFTW REDFT10 / REDFT01
  // definitions (ffts.f90:136-137)
  std::vector<double> tab(4 * (size_t)N);
  for (int t = 0; t < 4 * N; ++t) tab[t] = std::cos(pi * t / (2.0 * N));
  for (int q = 0; q < N; ++q)
    for (int n = 0; n < N; ++n) {
      c2[(size_t)q * N + n] = 2.0 / N * tab[((long)q * (2 * n + 1)) % (4 * N)];
      c3[(size_t)q * N + n] = 0.5 * (n == 0 ? 1.0 : 2.0 * tab[((long)n * (2 * q + 1)) % (4 * N)]);
    }
  dalloc(b.dct2, (size_t)N * N);
  dalloc(b.dct3, (size_t)N * N);
  h2d(b, b.dct2, c2.data(), c2.size() * sizeof(double));
  h2d(b, b.dct3, c3.data(), c3.size() * sizeof(double));
}

// Largest product of |c_j a2_j| over XLINE_WINDOW + 1 consecutive rows of
// a (3, n) tridiagonal system after the coefinv factorisation
// (coeft.f90:1338-1353, as factor_kernel does it on the GPU): the factor by
// which the windowed backward sweep damps its zero start.
static double window_decay(std::vector<double> a, int n) {
  auto A = [&](int r, int j) -> double & { return a[3 * (size_t)j + r - 1]; };
  for (int j = 2; j <= n; ++j) {
    double d = 1.0 / A(2, j - 1);
    A(2, j - 1) = d;
    A(1, j) = -A(1, j) * d;
    A(2, j) = A(2, j) + A(1, j) * A(3, j - 1);
  }
  A(2, n) = 1.0 / A(2, n);
  double worst = 0.0;
  for (int i = 1; i + XLINE_WINDOW <= n - 1; ++i) {
    double p = 1.0;
    for (int k = i; k <= i + XLINE_WINDOW; ++k) p *= std::fabs(A(3, k) * A(2, k));
    worst = std::max(worst, p);
  }
  return worst;
}

// Largest product of the factored multipliers |a(1,k)| (coeft.f90:1338-1353)
// over XLINE_LEAD consecutive rows: the factor by which a segment's forward
// elimination damps its zero start.
static double lead_decay(std::vector<double> a, int n) {
  auto A = [&](int r, int j) -> double & { return a[3 * (size_t)j + r - 1]; };
  for (int j = 2; j <= n; ++j) {
    double d = 1.0 / A(2, j - 1);
    A(2, j - 1) = d;
    A(1, j) = -A(1, j) * d;
    A(2, j) = A(2, j) + A(1, j) * A(3, j - 1);
  }
  double worst = 0.0;
  for (int i = 2; i + XLINE_LEAD - 1 <= n; ++i) {
    double p = 1.0;
    for (int k = i; k < i + XLINE_LEAD; ++k) p *= std::fabs(A(1, k));
    worst = std::max(worst, p);
  }
  return worst;
}

static void free_box(Box *b) {
  if (!b) return;
  for (int q = 0; q < 4; ++q)
    for (int f = 0; f < 2; ++f)
      if (b->gexec[q][f]) (void)hipGraphExecDestroy(b->gexec[q][f]);
  double *bufs[] = {b->coefbuf, b->resu, b->resv, b->resw, b->rhsu, b->rhsv, b->rhsw, b->wki1, b->wki2,
                    b->wki3, b->ya, b->yb, b->yc, b->rest, b->rest2, b->dct2, b->dct3, b->implu,
                    b->implv, b->um, b->u0, b->v0, b->gwork, b->dct_chirp, b->dct_bf, b->dct_tw4,
                    b->ytab, b->ztw, b->zrow, b->dct_ztw, b->rowmax, b->P0, b->S0, b->p0, b->um64, b->rm64,
                    b->diag, b->op64, b->rthcf, b->st, b->stcoef, b->stmean};
  for (double *p : bufs)
    if (p) (void)hipFree(p);
  if (b->diag_count) (void)hipFree(b->diag_count);
  float *fbufs[] = {b->coefbuf_f, b->ytab_f, b->ytabh_f, b->ztw_f, b->dct_chirp_f, b->dct_bf_f, b->dct_tw4_f, b->dct_ztw_f};
  for (float *p : fbufs)
    if (p) (void)hipFree(p);
  if (b->ens) (void)hipFree(b->ens);
  for (int f = 0; f < OTBL_NFIELDS; ++f)
    if (b->f[f]) (void)hipFree(b->f[f]);
  if (b->ds) (void)hipFree(b->ds);
  if (b->vmbits) (void)hipFree(b->vmbits);
  for (auto &p : b->pending) {
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  if (b->s2) (void)hipStreamDestroy(b->s2);
  for (hipEvent_t e : {b->ev_fork, b->ev_p1, b->ev_join})
    if (e) (void)hipEventDestroy(e);
  if (b->s) (void)hipStreamDestroy(b->s);
  delete b;
}

// With stats, substep 1 first adds a statistics sample of the state it
// starts from (rhsp.F90:72-102 at dostat, which bl_1.F90:92 sets before
// substep 1 and rhsp.F90:488 clears).
static void enqueue_substep(Box &b, int m, bool stats) {
  if (stats) k_statsp(b);
  k_substep_rhsp(b, m);
  k_outflow(b);
  k_pois(b, m, 0.0, true);
}

// the sampling rule of bl_1.F90:92, at the start of every step
static void begin_step(Box &b) {
  ++b.istep;
  if (b.st && b.avcrt > 0 && b.istep % b.avcrt == 0 && b.istep > b.stest) b.dostat = true;
}

// a sample requested for this substep (m = 1 only), then counted
static bool take_stats(Box &b, int m) { return m == 1 && b.dostat && b.st; }
static void stats_taken(Box &b) {
  b.dostat = false;
  ++b.ical; // rhsp.F90:92
}

static void run_substep(Box &b, int m) {
  const bool stats = take_stats(b, m);
  if (!b.use_graphs || b.kprof) {
    enqueue_substep(b, m, stats);
    if (stats) stats_taken(b);
    return;
  }
  // the substep swaps the velocity fields with their saved copies: one graph
  // per role assignment, and the host-side swap done again on replay
  const int fl = b.flip;
  hipGraphExec_t &ge = b.gexec[stats ? 0 : m][fl];
  if (!ge) {
    hipGraph_t g;
    OTBL_HIP(hipStreamBeginCapture(b.s, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_substep(b, m, stats);
    } catch (...) {
      hipGraph_t dummy = nullptr;
      if (hipStreamEndCapture(b.s, &dummy) == hipSuccess && dummy) (void)hipGraphDestroy(dummy);
      if (b.flip != fl) k_swap_saved(b);
      throw;
    }
    OTBL_HIP(hipStreamEndCapture(b.s, &g));
    OTBL_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    OTBL_HIP(hipGraphDestroy(g));
    if (b.flip != fl) k_swap_saved(b); // capture ran the host side only
  }
  OTBL_HIP(hipGraphLaunch(ge, b.s));
  k_swap_saved(b);
  if (stats) stats_taken(b);
}

} // namespace otbl

using namespace otbl;

struct otbl_ctx {
  Box *b;
};

#define API_BEGIN try {
#define API_END                                                                \
  return OTBL_OK;                                                              \
  }                                                                            \
  catch (const std::invalid_argument &e) {                                     \
    g_err = e.what();                                                          \
    return OTBL_EINVAL;                                                        \
  }                                                                            \
  catch (const std::bad_alloc &e) {                                            \
    g_err = "out of memory";                                                   \
    return OTBL_ENOMEM;                                                        \
  }                                                                            \
  catch (const std::exception &e) {                                            \
    g_err = e.what();                                                          \
    return OTBL_EHIP;                                                          \
  }

static Box &box_of(otbl_ctx *ctx) {
  if (!ctx || !ctx->b) throw std::invalid_argument("null otbl_ctx");
  return *ctx->b;
}

static void need(bool ok, const char *msg) {
  if (!ok) throw std::invalid_argument(msg);
}

// Slot in the FP64 mean-flow array of row r = (i-1) R + (j-1) of field id:
// p0 is [nx][ny] (the same order), um64 [ny+1][nx].
static size_t mean_slot(const Box &b, int id, size_t r) {
  if (id == OTBL_P) return r;
  const size_t R = b.ny + 1, i = r / R, j = r % R;
  return j * b.nx + i;
}

// FP32 copies of the factored FP64 coefficients and tables, the Poisson
// pivots and the FP64 mean-flow pencils of the mixed-precision substep.
static void setup_mixed(Box &b) {
  need(b.yL && b.dct_zfast && b.pchunk,
       "OTBL_MIXED needs the chunked y solver (ny <= 24*16) and a Poisson x transform on the register engine "
       "(nx - 1 from 129 to 4096)");
  need(b.xwin, "OTBL_MIXED needs x lines that the windowed sweep solves (nx <= 32 or decaying x systems)");
  size_t total = 0;
  for (const auto &sl : kSlots) total += (b.hc.*(sl.vec)).size() + 8;
  total += 7 * 24;
  dalloc(b.coefbuf_f, total);
  k_d2f(b, b.coefbuf, b.coefbuf_f, total);
  // the same offsets as the FP64 set
  const double *const *src = reinterpret_cast<const double *const *>(&b.dc);
  const float **dst = reinterpret_cast<const float **>(&b.dcf);
  static_assert(sizeof(DevCoefT<double>) == sizeof(DevCoefT<float>), "coefficient pointer sets");
  for (size_t q = 0; q < sizeof(DevCoef) / sizeof(void *); ++q) dst[q] = b.coefbuf_f + (src[q] - b.coefbuf);
  dalloc(b.ytab_f, YT_COUNT * ytab_size(b.yL));
  k_d2f(b, b.ytab, b.ytab_f, YT_COUNT * ytab_size(b.yL));
  // the half-wave rk_final addresses a field by 32-bit byte offsets: fields
  // of 2^30 FP32 values or more take the full-wave kernel
  const char *yh = std::getenv("OTBL_Y_HALF");
  if (b.yL % 2 == 0 && !(yh && yh[0] == '0') && b.nu < (size_t(1) << 30)) {
    dalloc(b.ytabh_f, YT_COUNT * ytabh_size(b.yL / 2));
    k_ytab_halve_fixed(b);
  }
  // the z twiddle table exists only for the register-resident transforms
  // (the LDS Stockham kernels build their twiddles in LDS)
  if (b.zfast) {
    dalloc(b.ztw_f, 2 * (size_t)b.nz);
    k_d2f(b, b.ztw, b.ztw_f, 2 * (size_t)b.nz);
  }
  const size_t N = b.nx - 1, M = b.dct_M;
  dalloc(b.dct_chirp_f, 2 * N);
  dalloc(b.dct_tw4_f, 2 * N);
  dalloc(b.dct_bf_f, 2 * M);
  dalloc(b.dct_ztw_f, 2 * M);
  k_d2f(b, b.dct_chirp, b.dct_chirp_f, 2 * N);
  k_d2f(b, b.dct_tw4, b.dct_tw4_f, 2 * N);
  k_d2f(b, b.dct_bf, b.dct_bf_f, 2 * M);
  k_d2f(b, b.dct_ztw, b.dct_ztw_f, 2 * M);
  dalloc(b.P0, 2 * (size_t)(b.ny - 1) * N);
  dalloc(b.S0, 2 * (size_t)(b.ny - 1) * N);
  dalloc(b.p0, (size_t)b.nx * b.ny);
  dalloc(b.um64, (size_t)b.nx * (b.ny + 1));
  dalloc(b.rm64, (size_t)b.nx * (b.ny + 1));
}

extern "C" {

const char *otbl_last_error(void) { return g_err.c_str(); }

int otbl_create(const otbl_params *prm, const double *y, otbl_ctx **out) {
  Box *b = nullptr;
  API_BEGIN
  need(prm && y && out, "otbl_create: null argument");
  *out = nullptr;
  need(prm->nx >= 8, "nx must be >= 8");
  need(prm->ny >= 24, "ny must be >= 24 (the free-stream average spans rows ny-19..ny-4)");
  need(prm->nz >= 6 && prm->nz % 2 == 0, "nz must be even and >= 6");
  need(prm->xout >= 2 && prm->xout <= prm->nx - 1, "xout must lie in 2..nx-1 (bl_1.F90:287-297)");
  need(prm->re > 0 && prm->ax > 0 && prm->az > 0 && prm->cfl > 0, "re, ax, az, cfl must be > 0");
  need(prm->precision == OTBL_FP64 || prm->precision == OTBL_MIXED, "precision must be OTBL_FP64 or OTBL_MIXED");
  need(prm->dt_mode == OTBL_DT_EVERY_STEP || prm->dt_mode == OTBL_DT_FROZEN, "bad dt_mode");
  b = new Box();
  b->nx = prm->nx; b->ny = prm->ny; b->nz = prm->nz;
  b->nz1 = 2 * (prm->nz / 3);
  b->nz2 = b->nz1 / 2 - 1;
  b->mz = b->nz2 + 1;
  b->xout = prm->xout;
  b->precision = prm->precision;
  b->esz = prm->precision == OTBL_MIXED ? 4 : 8;
  b->dt_mode = prm->dt_mode;
  b->re = prm->re; b->ax = prm->ax; b->az = prm->az; b->cfl = prm->cfl;
  {
    int fac[32];
    need(small_factors(b->nz, fac) > 0, "nz must factor into 2, 3, 5 and 7");
  }
  b->zplan = make_plan(b->nz);
  need(k_zfft_shm_bytes(b->nz, 4) <= 160 * 1024, "nz too large for the FP64 LDS transform");
  b->pu = (size_t)b->nz1 * (b->ny + 1);
  b->pv = (size_t)b->nz1 * b->ny;
  b->nu = b->pu * b->nx;
  b->nv = b->pv * b->nx;
  need(b->nu < (size_t(1) << 32), "a field must have fewer than 2^32 elements");
  OTBL_HIP(hipGetDevice(&b->dev));
  (void)scratch_of(b->dev); // the no-allocation entry point's scratch on this device
  OTBL_HIP(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking));
  build_coef(b->hc, b->nx, b->ny, b->nz, b->re, b->ax, b->az, b->cfl, y);
  upload_coef(*b);
  for (int f = 0; f < OTBL_NFIELDS; ++f)
    falloc(*b, b->f[f], (f == OTBL_V || f == OTBL_P || f == OTBL_RHSVPA) ? b->nv : b->nu);
  falloc(*b, b->resu, b->nu); falloc(*b, b->resv, b->nv); falloc(*b, b->resw, b->nu);
  falloc(*b, b->rhsu, b->nu); falloc(*b, b->rhsv, b->nv); falloc(*b, b->rhsw, b->nu);
  falloc(*b, b->wki1, b->nu); falloc(*b, b->wki2, b->nv); falloc(*b, b->wki3, b->nu);
  falloc(*b, b->ya, b->nu); falloc(*b, b->yb, b->nu); falloc(*b, b->yc, b->nu);
  b->pld = (b->nx - 1 + 31) / 32 * 32;
  b->phs = (b->pv + 31) / 32 * 32;
  {
    // the pencils (and the FP64 solve's scratch in the same layout) live in
    // rest2 / rest: (ny-1) nz1 lines of pld values, phi nx planes of phs
    // values, or a v-type field
    const size_t pen = std::max({b->nv, (size_t)(b->ny - 1) * b->nz1 * b->pld, (size_t)b->nx * b->phs});
    falloc(*b, b->rest, pen); falloc(*b, b->rest2, pen);
  }
  dalloc(b->implu, 3 * (size_t)(b->ny + 2) + 8);
  dalloc(b->implv, 3 * (size_t)(b->ny + 2) + 8);
  dalloc(b->um, b->ny + 2); dalloc(b->u0, b->ny + 2); dalloc(b->v0, b->ny + 2);
  dalloc(b->gwork, (size_t)b->nz1 * (b->ny + 1) + 16 * (size_t)(b->ny + 2));
  OTBL_HIP(hipMalloc((void **)&b->ds, sizeof(DevScalars)));
  OTBL_HIP(hipMemset(b->ds, 0, sizeof(DevScalars)));
  OTBL_HIP(hipDeviceSynchronize());
  dalloc(b->vmbits, b->ny + 1);
  {
    // FFT-based Poisson x transforms: on the register engine when the
    // power-of-two convolution length is instantiated, else on the LDS
    // Stockham kernel when its buffers fit; dense matrices otherwise
    const int N = b->nx - 1;
    const bool zf = zdct_length(2 * N - 1) != 0 && std::getenv("OTBL_Z_STOCKHAM") == nullptr;
    const int M = good_length(2 * N - 1);
    b->dct_fft = std::getenv("OTBL_DCT_GEMM") == nullptr && (zf || k_dct_shm_bytes(M) <= 160 * 1024);
  }
  if (b->dct_fft)
    build_dct_fft(*b);
  else
    build_dct(*b);
  k_setup_attrs(*b);
  // register-resident z transforms for the lengths they are instantiated
  // for (OTBL_Z_STOCKHAM keeps the LDS Stockham kernels)
  b->zfast = zfft_supported(b->nz) && std::getenv("OTBL_Z_STOCKHAM") == nullptr;
  b->zpair = std::getenv("OTBL_Z_NOPAIR") ? 0 : (std::getenv("OTBL_ZP_OCC2") ? 2 : 1);
  b->ytail = std::getenv("OTBL_Y_NOTAIL") == nullptr;
  dalloc(b->rowmax, 3 * (size_t)(b->ny + 1) * b->nx);
  if (b->zfast) {
    dalloc(b->ztw, 2 * (size_t)b->nz);
    dalloc(b->zrow, (size_t)b->nz1);
    OTBL_HIP(hipMemset(b->zrow, 0, (size_t)b->nz1 * sizeof(double)));
    OTBL_HIP(hipDeviceSynchronize());
    kz_twiddles(*b, b->ztw, b->nz);
  }
  if (b->dct_zfast) {
    dalloc(b->dct_ztw, 2 * (size_t)b->dct_M);
    kz_twiddles(*b, b->dct_ztw, b->dct_M);
  }
  k_factor_coef(*b, b->coefbuf);
  // windowed x lines when the lines are long enough and every x system
  // damps the window's zero start below round-off
  {
    const HostCoef &h = b->hc;
    double worst = 0.0;
    for (const std::vector<double> *a : {&h.inxu, &h.inxv, &h.dcxu, &h.dcxv, &h.vixu, &h.vixv})
      worst = std::max(worst, window_decay(*a, b->nx));
    b->xwin = b->nx > 2 * XLINE_WINDOW && worst < 1e-17 && std::getenv("OTBL_X_MARCH") == nullptr;
    // the mixed substep has no marching fallback: short lines, which the
    // windowed kernel solves exactly, or any decaying x systems
    if (b->precision == OTBL_MIXED) b->xwin = b->nx <= XLINE_EXACT || worst < 1e-17;
    // segmented lines when the forward recurrences forget their start as
    // well: about 6e5 threads in flight, at most 8 segments of at least one
    // group of XLINE_LEAD planes.  A segment marches its lead-in and the
    // backward window past its end again; on short lines the parallelism
    // pays for that (interpxx + difvisxx of a 256 x 96 x 64 box: 1.46 ms on
    // one segment, 0.44 ms on eight; 1536 x 315 x 1024 keeps three)
    double lead = 0.0;
    for (const std::vector<double> *a : {&h.inxu, &h.inxv, &h.dcxu, &h.dcxv, &h.vixu, &h.vixv})
      lead = std::max(lead, lead_decay(*a, b->nx));
    const size_t lines = b->pu;
    int nseg = (int)std::min<size_t>(8, std::max<size_t>(1, (600000 + lines / 2) / lines));
    nseg = std::max(1, std::min(nseg, b->nx / XLINE_LEAD));
    // OTBL_XSEG=n forces n segments (at most one per group of XLINE_LEAD planes)
    if (const char *e = std::getenv("OTBL_XSEG"))
      nseg = std::max(1, std::min(std::atoi(e), (b->nx + XLINE_LEAD - 1) / XLINE_LEAD));
    b->xseg = (b->xwin && lead < 1e-17) ? nseg : 1;
    // the most segments the autotuner may try: at most 8, each at least one
    // group of XLINE_LEAD planes
    const int smax = std::max(1, std::min(8, b->nx / XLINE_LEAD));
    b->xseg_max = (b->xwin && lead < 1e-17) ? std::max(smax, b->xseg) : 1;
  }
  // chunked y-line solver unless the columns are longer than its
  // instantiations (or OTBL_Y_MARCH asks for one thread per column)
  b->yL = std::getenv("OTBL_Y_MARCH") ? 0 : ytab_rows(b->ny);
  b->pchunk = std::getenv("OTBL_Y_MARCH") ? 0 : kp_chunk_rows(b->ny);
  if (b->yL) {
    dalloc(b->ytab, YT_COUNT * ytab_size(b->yL));
    k_ytab_fixed(*b);
  }
  // launch limits of the partitioned Poisson solve, checked here so that a
  // step fails only before any state changes
  if (b->pchunk)
    need((size_t)(2 * b->pchunk + 1) * b->nz1 * (b->nx - 1) * b->esz < (size_t(1) << 32),
         "Poisson pencil rows too long for the partitioned y solve's 32-bit offsets");
  if (b->precision == OTBL_MIXED) setup_mixed(*b);
  // the second stream of rhsp's plane loop 2, on the paths that keep ya and
  // yb free of scratch use (windowed x sweep, chunked y solver, register z
  // engine)
  if (b->xwin && b->yL && b->zfast && b->nx > 2 && std::getenv("OTBL_SERIAL_LOOPS") == nullptr) {
    OTBL_HIP(hipStreamCreateWithFlags(&b->s2, hipStreamNonBlocking));
    OTBL_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    OTBL_HIP(hipEventCreateWithFlags(&b->ev_p1, hipEventDisableTiming));
    OTBL_HIP(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
  }
  OTBL_HIP(hipStreamSynchronize(b->s));
  *out = new otbl_ctx{b};
  b = nullptr;
  return OTBL_OK;
  }
  catch (const std::invalid_argument &e) {
    free_box(b);
    g_err = e.what();
    return OTBL_EINVAL;
  }
  catch (const std::exception &e) {
    free_box(b);
    g_err = e.what();
    return OTBL_EHIP;
  }
}

int otbl_destroy(otbl_ctx *ctx) {
  if (!ctx) return OTBL_OK;
  if (ctx->b) {
    (void)hipStreamSynchronize(ctx->b->s);
    free_box(ctx->b);
  }
  delete ctx;
  return OTBL_OK;
}

size_t otbl_field_len(const otbl_ctx *ctx, int id) {
  if (!ctx || !ctx->b || id < 0 || id >= OTBL_NFIELDS) return 0;
  const Box &b = *ctx->b;
  return (id == OTBL_V || id == OTBL_P || id == OTBL_RHSVPA) ? b.nv : b.nu;
}

int otbl_set_field(otbl_ctx *ctx, int id, const double *host) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(id >= 0 && id < OTBL_NFIELDS && host, "otbl_set_field: bad id or null pointer");
  OTBL_HIP(hipStreamSynchronize(b.s));
  const size_t n = otbl_field_len(ctx, id);
  if (b.esz == 8) {
    h2d(b, b.f[id], host, n * sizeof(double));
  } else {
    // staged through a work field (nu FP32 elements: room for nu / 2
    // doubles at a time), rounded to FP32
    const size_t chunk = b.nu / 2;
    for (size_t o = 0; o < n; o += chunk) {
      const size_t m = std::min(chunk, n - o);
      h2d(b, b.ya, host + o, m * sizeof(double));
      k_d2f(b, b.ya, fld<float>(b.f[id]) + o, m);
      OTBL_HIP(hipStreamSynchronize(b.s));
    }
    if (id == OTBL_P || id == OTBL_U) {
      // the mean-flow pressure ([nx][ny]) and u ([ny+1][nx]) (real parts of
      // mode 0) are kept in FP64
      std::vector<double> m(n / b.nz1);
      for (size_t r = 0; r < m.size(); ++r) m[mean_slot(b, id, r)] = host[r * b.nz1];
      h2d(b, id == OTBL_P ? b.p0 : b.um64, m.data(), m.size() * sizeof(double));
    }
  }
  API_END
}

int otbl_get_field(otbl_ctx *ctx, int id, double *host) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(id >= 0 && id < OTBL_NFIELDS && host, "otbl_get_field: bad id or null pointer");
  OTBL_HIP(hipStreamSynchronize(b.s));
  const size_t n = otbl_field_len(ctx, id);
  if (b.esz == 8) {
    d2h(b, host, b.f[id], n * sizeof(double));
  } else {
    const size_t chunk = b.nu / 2;
    for (size_t o = 0; o < n; o += chunk) {
      const size_t m = std::min(chunk, n - o);
      k_f2d(b, fld<float>(b.f[id]) + o, b.ya, m);
      OTBL_HIP(hipStreamSynchronize(b.s));
      d2h(b, host + o, b.ya, m * sizeof(double));
    }
    if (id == OTBL_P || id == OTBL_U) {
      std::vector<double> m(n / b.nz1);
      d2h(b, m.data(), id == OTBL_P ? b.p0 : b.um64, m.size() * sizeof(double));
      for (size_t r = 0; r < m.size(); ++r) host[r * b.nz1] = m[mean_slot(b, id, r)];
    }
  }
  API_END
}

int otbl_set_inflow(otbl_ctx *ctx, const double *u0, const double *v0, const double *um) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(u0 && v0 && um, "otbl_set_inflow: null pointer");
  OTBL_HIP(hipStreamSynchronize(b.s));
  double rthin = 0, din = 0;
  inlet_setup(b.hc, u0, &rthin, &din);
  h2d(b, b.u0, u0, (b.ny + 1) * sizeof(double));
  h2d(b, b.v0, v0, b.ny * sizeof(double));
  h2d(b, b.um, um, (b.ny + 1) * sizeof(double));
  h2d(b, (void *)b.dc.vmagic, b.hc.vmagic.data(), b.hc.vmagic.size() * sizeof(double));
  if (b.dcf.vmagic) {
    k_d2f(b, b.dc.vmagic, (float *)b.dcf.vmagic, b.hc.vmagic.size());
    OTBL_HIP(hipStreamSynchronize(b.s));
  }
  put_ds(b, offsetof(DevScalars, rthin), &rthin, sizeof(double));
  put_ds(b, offsetof(DevScalars, din), &din, sizeof(double));
  API_END
}

int otbl_get_um(otbl_ctx *ctx, double *um) {
  API_BEGIN
  Box &b = box_of(ctx);
  OTBL_HIP(hipStreamSynchronize(b.s));
  d2h(b, um, b.um, (b.ny + 1) * sizeof(double));
  API_END
}

int otbl_get_genflu(otbl_ctx *ctx, double *o) {
  API_BEGIN
  Box &b = box_of(ctx);
  DevScalars d;
  get_ds(b, 0, &d, sizeof d);
  o[0] = d.rthin; o[1] = d.din; o[2] = d.rthout; o[3] = d.dout; o[4] = d.gamma1; o[5] = d.gamma2;
  API_END
}

int otbl_get_ener(otbl_ctx *ctx, double *e) {
  API_BEGIN
  Box &b = box_of(ctx);
  DevScalars d;
  get_ds(b, 0, &d, sizeof d);
  for (int q = 0; q < 15; ++q) e[q] = d.ener[q + 1];
  API_END
}

// graphs bake kernel arguments in: drop them when the launch set changes
static void drop_graphs(Box &b) {
  ++b.graph_gen;
  for (int q = 0; q < 4; ++q)
    for (int f = 0; f < 2; ++f)
      if (b.gexec[q][f]) {
        (void)hipGraphExecDestroy(b.gexec[q][f]);
        b.gexec[q][f] = nullptr;
      }
}

int otbl_diag_enable(otbl_ctx *ctx, int capacity) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(capacity >= 0, "otbl_diag_enable: capacity must be >= 0");
  OTBL_HIP(hipStreamSynchronize(b.s));
  drop_graphs(b);
  if (b.diag) (void)hipFree(b.diag);
  if (b.diag_count) (void)hipFree(b.diag_count);
  b.diag = nullptr;
  b.diag_count = nullptr;
  b.diag_cap = 0;
  b.diag_read = 0;
  if (capacity > 0) {
    dalloc(b.diag, (size_t)capacity * OTBL_DIAG_LEN);
    dalloc(b.diag_count, 1);
    b.diag_cap = capacity;
  }
  API_END
}

int otbl_diag_read(otbl_ctx *ctx, double *out, int max_records, int *n_records, long long *dropped) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(b.diag != nullptr, "otbl_diag_read: diagnostics are off (otbl_diag_enable)");
  need(max_records >= 0 && (out != nullptr || max_records == 0), "otbl_diag_read: bad output buffer");
  OTBL_HIP(hipStreamSynchronize(b.s));
  unsigned long long n = 0;
  d2h(b, &n, b.diag_count, sizeof n);
  const unsigned long long cap = (unsigned long long)b.diag_cap;
  long long lost = 0;
  if (n - b.diag_read > cap) {
    lost = (long long)(n - cap - b.diag_read);
    b.diag_read = n - cap;
  }
  const unsigned long long take = std::min<unsigned long long>(n - b.diag_read, (unsigned long long)max_records);
  std::vector<double> ring((size_t)b.diag_cap * OTBL_DIAG_LEN);
  if (take) d2h(b, ring.data(), b.diag, ring.size() * sizeof(double));
  for (unsigned long long r = 0; r < take; ++r) {
    const size_t slot = (size_t)((b.diag_read + r) % cap);
    std::copy(ring.begin() + slot * OTBL_DIAG_LEN, ring.begin() + (slot + 1) * OTBL_DIAG_LEN,
              out + r * OTBL_DIAG_LEN);
  }
  b.diag_read += take;
  if (n_records) *n_records = (int)take;
  if (dropped) *dropped = lost;
  API_END
}

int otbl_retheta_cf(otbl_ctx *ctx, double *rth, double *cf) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(rth && cf, "otbl_retheta_cf: null pointer");
  if (!b.rthcf) {
    OTBL_HIP(hipStreamSynchronize(b.s));
    dalloc(b.rthcf, 2 * (size_t)b.nx);
  }
  k_retheta_cf(b, b.rthcf);
  std::vector<double> h(2 * (size_t)b.nx);
  d2h(b, h.data(), b.rthcf, h.size() * sizeof(double));
  std::copy(h.begin(), h.begin() + b.nx, rth);
  std::copy(h.begin() + b.nx, h.end(), cf);
  API_END
}

int otbl_set_time(otbl_ctx *ctx, double tiempo, double timeinit, double dt) {
  API_BEGIN
  Box &b = box_of(ctx);
  OTBL_HIP(hipStreamSynchronize(b.s));
  int valid = dt > 0.0 ? 1 : 0;
  put_ds(b, offsetof(DevScalars, tiempo), &tiempo, sizeof(double));
  put_ds(b, offsetof(DevScalars, timeinit), &timeinit, sizeof(double));
  put_ds(b, offsetof(DevScalars, dt), &dt, sizeof(double));
  put_ds(b, offsetof(DevScalars, dt_valid), &valid, sizeof(int));
  API_END
}

int otbl_get_time(otbl_ctx *ctx, double *tiempo, double *dt) {
  API_BEGIN
  Box &b = box_of(ctx);
  DevScalars d;
  get_ds(b, 0, &d, sizeof d);
  if (tiempo) *tiempo = d.tiempo;
  if (dt) *dt = d.dt;
  API_END
}

int otbl_get_coef(otbl_ctx *ctx, const char *name, double *out, int *rows, int *cols) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(name && rows && cols, "otbl_get_coef: null argument");
  OTBL_HIP(hipStreamSynchronize(b.s));
  for (const auto &sl : kSlots) {
    if (strcmp(sl.name, name)) continue;
    const auto &hv = b.hc.*(sl.vec);
    std::vector<double> dv(hv.size());
    d2h(b, dv.data(), b.dc.*(sl.dev), dv.size() * sizeof(double));
    const int r = sl.rows, first = sl.first, n = slot_cols(b, sl.cols);
    *rows = r;
    *cols = n;
    if (out)
      for (int q = 0; q < r * n; ++q) out[q] = dv[(size_t)r * first + q];
    return OTBL_OK;
  }
  // the stats-only schemes (8, 1:n), LU-factored, from the host set
  const std::vector<double> *fd = !strcmp(name, "fd_dvdy") ? &b.hc.fd_dvdy
                                  : !strcmp(name, "fd_iy")  ? &b.hc.fd_iy
                                  : !strcmp(name, "fd_iyp") ? &b.hc.fd_iyp
                                  : !strcmp(name, "fd_dx")  ? &b.hc.fd_dx
                                  : !strcmp(name, "fd_ix")  ? &b.hc.fd_ix
                                                            : nullptr;
  if (fd) {
    *rows = 8;
    *cols = (int)(fd->size() / 8) - 1;
    if (out) std::copy(fd->begin() + 8, fd->end(), out);
    return OTBL_OK;
  }
  const size_t n = fixed_len(name);
  const double *src = nullptr;
  int r = 4, c = 4;
  if (!strcmp(name, "inbx")) src = b.dc.inbx;
  else if (!strcmp(name, "inby")) src = b.dc.inby;
  else if (!strcmp(name, "dcbx")) { src = b.dc.dcbx; r = 5; }
  else if (!strcmp(name, "dcby")) src = b.dc.dcby;
  else if (!strcmp(name, "ccon")) { src = b.dc.ccon; r = 2; }
  else if (!strcmp(name, "cofvbx")) src = b.dc.cofvbx;
  else if (!strcmp(name, "cofvby")) src = b.dc.cofvby;
  need(src != nullptr, "otbl_get_coef: unknown coefficient name");
  *rows = r;
  *cols = c;
  if (out) d2h(b, out, src, n * sizeof(double));
  API_END
}

int otbl_rhsp(otbl_ctx *ctx, int m) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(m >= 1 && m <= 3, "substep m must be 1..3");
  const bool stats = take_stats(b, m);
  if (stats) k_statsp(b);
  k_substep_rhsp(b, m);
  if (stats) stats_taken(b);
  API_END
}

int otbl_outflow_correction(otbl_ctx *ctx) {
  API_BEGIN
  k_outflow(box_of(ctx));
  API_END
}

int otbl_pois(otbl_ctx *ctx, double varstep) {
  API_BEGIN
  k_pois(box_of(ctx), 0, varstep, false);
  API_END
}

int otbl_pois_m(otbl_ctx *ctx, int m) {
  API_BEGIN
  need(m >= 1 && m <= 3, "substep m must be 1..3");
  k_pois(box_of(ctx), m, 0.0, true);
  API_END
}

int otbl_substep(otbl_ctx *ctx, int m) {
  API_BEGIN
  need(m >= 1 && m <= 3, "substep m must be 1..3");
  run_substep(box_of(ctx), m);
  API_END
}

int otbl_step(otbl_ctx *ctx, int nsteps) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(nsteps >= 0, "nsteps must be >= 0");
  for (int s = 0; s < nsteps; ++s) {
    begin_step(b);
    for (int m = 1; m <= 3; ++m) run_substep(b, m);
    k_end_step(b);
  }
  API_END
}

int otbl_use_graphs(otbl_ctx *ctx, int on) {
  API_BEGIN
  box_of(ctx).use_graphs = on != 0;
  API_END
}

int otbl_sync(otbl_ctx *ctx) {
  API_BEGIN
  Box &b = box_of(ctx);
  OTBL_HIP(hipStreamSynchronize(b.s));
  drain_kprof(b);
  API_END
}

int otbl_interpxx(otbl_ctx *ctx, const double *gen, double *geni, int R, int which_u) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(R == b.ny || R == b.ny + 1, "R must be ny or ny+1");
  k_interpxx(b, gen, geni, R, which_u);
  API_END
}

int otbl_difvisxx(otbl_ctx *ctx, const double *prod, const double *vel, double *out, int R, int which_u) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(R == b.ny || R == b.ny + 1, "R must be ny or ny+1");
  k_difvisxx(b, prod, vel, out, R, which_u, 0);
  API_END
}

int otbl_interpyy(otbl_ctx *ctx, const double *gen, double *geni, int loguv, int nplanes) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(nplanes >= 1 && nplanes <= b.nx, "nplanes out of range");
  k_interpyy(b, gen, geni, loguv ? 1 : 0, 1, nplanes, loguv ? b.pv : b.pu, b.pu);
  API_END
}

int otbl_differzy(otbl_ctx *ctx, const double *geni, double *res, int which, int nplanes) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(nplanes >= 1 && nplanes <= b.nx, "nplanes out of range");
  k_differzy(b, geni, res, which ? 1 : 0, nplanes, b.pu, b.pu);
  API_END
}

int otbl_vistyy(otbl_ctx *ctx, const double *u, double *res, int R, int nplanes) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(R == b.ny || R == b.ny + 1, "R must be ny or ny+1");
  need(nplanes >= 1 && nplanes <= b.nx, "nplanes out of range");
  k_vistyy(b, u, res, R, nplanes);
  API_END
}

int otbl_implzy(otbl_ctx *ctx, double *u, const double *res, int R, double rkk, int nplanes) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(R == b.ny || R == b.ny + 1, "R must be ny or ny+1");
  need(nplanes >= 1 && nplanes <= b.nx, "nplanes out of range");
  k_implzy_op(b, u, res, R, rkk, nplanes);
  API_END
}

int otbl_fourxz(otbl_ctx *ctx, double *fou, double *phys, int iopt, int nrows) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(nrows >= 1, "nrows must be >= 1");
  k_fourxz(b, fou, phys, iopt, nrows);
  API_END
}

int otbl_cosftx(otbl_ctx *ctx, const double *in, double *out, int iopt) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(in != out, "otbl_cosftx works out of place");
  k_cosftx(b, in, out, iopt);
  API_END
}

int otbl_cosft_lines(otbl_ctx *ctx, double *lines, int nlines, int iopt) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(lines && nlines >= 1, "otbl_cosft_lines: null buffer or no lines");
  k_cosft_lines(b, lines, nlines, iopt);
  API_END
}

int otbl_soltriy(otbl_ctx *ctx, const double *u, double *us, int ii, double *a, int ny1) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(u && us && a && ny1 >= 1, "otbl_soltriy: null argument or ny1 < 1");
  need(ii >= 0 && ii <= b.nz2, "otbl_soltriy: column outside 0..nz2");
  double *da = nullptr;
  const size_t bytes = 3 * (size_t)(ny1 + 1) * sizeof(double);
  OTBL_HIP(hipStreamSynchronize(b.s));
  OTBL_HIP(hipMalloc((void **)&da, bytes));
  try {
    // the host array is (3, 1:ny1), stored from row 1: column j at 3 (j - 1)
    h2d(b, da + 3, a, 3 * (size_t)ny1 * sizeof(double));
    k_soltriy(b, u, us, ii, da, ny1, b.nz2 + 1);
    d2h(b, a, da + 3, 3 * (size_t)ny1 * sizeof(double));
  } catch (...) {
    (void)hipFree(da);
    throw;
  }
  OTBL_HIP(hipFree(da));
  API_END
}

int otbl_soltriy_noalloc(otbl_ctx *ctx, const double *u, double *us, int ii, double *a, int ny1) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(u && us && a && ny1 >= 1, "otbl_soltriy_noalloc: null argument or ny1 < 1");
  need(ny1 <= kNoAllocRows, "otbl_soltriy_noalloc: ny1 above 8192, the size of its scratch");
  need(ii >= 0 && ii <= b.nz2, "otbl_soltriy_noalloc: column outside 0..nz2");
  // every element the kernel reads (rows 1..3 of columns 1..ny1) is written
  // by the copy below, so the scratch needs no clearing between calls
  double *da = scratch_of(b.dev);
  OTBL_HIP(hipStreamSynchronize(b.s));
  h2d(b, da + 3, a, 3 * (size_t)ny1 * sizeof(double));
  k_soltriy(b, u, us, ii, da, ny1, b.nz2 + 1);
  d2h(b, a, da + 3, 3 * (size_t)ny1 * sizeof(double));
  API_END
}

int otbl_coefinv(otbl_ctx *ctx, double *a, int npter1, int neq) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(a && neq >= npter1, "otbl_coefinv: null array or neq < npter1");
  const int nrow = neq - npter1 + 1;
  double *da = nullptr;
  OTBL_HIP(hipStreamSynchronize(b.s));
  OTBL_HIP(hipMalloc((void **)&da, 3 * (size_t)(nrow + 1) * sizeof(double)));
  try {
    h2d(b, da + 3, a, 3 * (size_t)nrow * sizeof(double));
    k_coefinv(b, da, nrow);
    d2h(b, a, da + 3, 3 * (size_t)nrow * sizeof(double));
  } catch (...) {
    (void)hipFree(da);
    throw;
  }
  OTBL_HIP(hipFree(da));
  API_END
}

int otbl_soltrix(otbl_ctx *ctx, const double *u, int nstr, int nlen, int neq, double *us, int npter1, int npter2,
                 double *a, int *iflag) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(u && us && a && iflag, "otbl_soltrix: null argument");
  need(neq >= 1 && nlen >= 0 && npter1 >= 1 && npter2 >= 1 && npter1 + neq - 1 <= nstr,
       "otbl_soltrix: rows npter1 .. npter1 + neq - 1 must lie in 1 .. nstr, neq >= 1, nlen >= 0");
  // a(1:3, npter1:nfin) at a + 3 (npter1 - 1); columns npter2 .. of u, us
  double *ap = a + 3 * (size_t)(npter1 - 1);
  const size_t c0 = (size_t)(npter2 - 1) * nstr + (size_t)(npter1 - 1), ncol = (size_t)nlen * nstr;
  double *da = nullptr, *du = nullptr;
  OTBL_HIP(hipStreamSynchronize(b.s));
  OTBL_HIP(hipMalloc((void **)&da, 3 * (size_t)(neq + 1) * sizeof(double)));
  try {
    if (nlen > 0) OTBL_HIP(hipMalloc((void **)&du, 2 * ncol * sizeof(double)));
    h2d(b, da + 3, ap, 3 * (size_t)neq * sizeof(double));
    if (*iflag == 0) k_coefinv(b, da, neq);
    if (nlen > 0) {
      // rows npter1 .. nfin of each column: u's into du, us's into du + ncol
      // (us keeps its other rows)
      const size_t pitch = (size_t)nstr * sizeof(double), w = (size_t)neq * sizeof(double);
      OTBL_HIP(hipMemcpy2DAsync(du, pitch, u + c0, pitch, w, nlen, hipMemcpyHostToDevice, b.s));
      k_soltrix(b, du, du + ncol, da + 3, nstr, nlen, neq);
      OTBL_HIP(hipMemcpy2DAsync(us + c0, pitch, du + ncol, pitch, w, nlen, hipMemcpyDeviceToHost, b.s));
    }
    if (*iflag == 0) d2h(b, ap, da + 3, 3 * (size_t)neq * sizeof(double));
    OTBL_HIP(hipStreamSynchronize(b.s));
  } catch (...) {
    (void)hipStreamSynchronize(b.s);
    (void)hipFree(da);
    if (du) (void)hipFree(du);
    throw;
  }
  OTBL_HIP(hipFree(da));
  if (du) OTBL_HIP(hipFree(du));
  if (*iflag == 0) *iflag = 1;
  API_END
}

int otbl_coefb(int n, int flagv, double a, double b, double c, double d, double *bcc) {
  API_BEGIN
  need(bcc != nullptr, "otbl_coefb: null bcc");
  double w[5];
  coefb_weights(n, flagv, a, b, c, d, w);
  for (int q = 0; q < n; ++q) bcc[q] = w[q];
  API_END
}

int otbl_dev_alloc(size_t bytes, void **ptr) {
  API_BEGIN
  need(ptr != nullptr, "null pointer");
  OTBL_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  API_END
}

int otbl_dev_free(void *ptr) {
  API_BEGIN
  if (ptr) OTBL_HIP(hipFree(ptr));
  API_END
}

int otbl_h2d(void *dst, const void *src, size_t bytes) {
  API_BEGIN
  OTBL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  OTBL_HIP(hipDeviceSynchronize());
  API_END
}

int otbl_d2h(void *dst, const void *src, size_t bytes) {
  API_BEGIN
  OTBL_HIP(hipDeviceSynchronize());
  OTBL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  API_END
}

int otbl_d2d(void *dst, const void *src, size_t bytes) {
  API_BEGIN
  OTBL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  OTBL_HIP(hipDeviceSynchronize());
  API_END
}

int otbl_get_paths(otbl_ctx *ctx, char *buf, size_t len) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(buf && len > 0, "otbl_get_paths: null buffer");
  char tmp[256];
  std::snprintf(tmp, sizeof tmp,
                "y:%s%d%s z:%s%d dct:%s%d x:%s xseg:%d pois:%s%d prun:%d tuned:%d streams:%d tail:%d",
                b.yL ? "chunk" : "march", b.yL, (b.ytabh_f && b.yhalf) ? "h" : "", b.zfast ? "engine" : "stockham", b.nz,
                b.dct_zfast ? "engine" : (b.dct_fft ? "stockham" : "gemm"), b.dct_M, b.xwin ? "window" : "march", b.xseg,
                b.pchunk ? "chunk" : "march", b.pchunk, b.prun > 0 ? b.prun : (b.esz == 4 ? 4 : 2), b.tuned ? 1 : 0,
                b.s2 ? 2 : 1, b.tail() ? 1 : 0);
  std::snprintf(buf, len, "%s", tmp);
  API_END
}

int otbl_autotune(otbl_ctx *ctx, int reps, char *report, size_t len) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(reps >= 1 && reps <= 1000, "otbl_autotune: reps must be 1..1000");
  need(!b.kprof, "otbl_autotune: turn kprof off first");
  OTBL_HIP(hipStreamSynchronize(b.s));
  const std::string rep = k_autotune(b, reps);
  OTBL_HIP(hipStreamSynchronize(b.s));
  drop_graphs(b); // captured graphs hold the old launch shapes
  if (report && len > 0) std::snprintf(report, len, "%s", rep.c_str());
  API_END
}

int otbl_kprof(otbl_ctx *ctx, int on) {
  API_BEGIN
  Box &b = box_of(ctx);
  drain_kprof(b);
  b.kprof = on != 0;
  if (on) b.kstat.clear();
  API_END
}

int otbl_kprof_report(otbl_ctx *ctx, char *buf, size_t len) {
  API_BEGIN
  Box &b = box_of(ctx);
  drain_kprof(b);
  std::string s;
  char line[256];
  for (auto &kv : b.kstat) {
    snprintf(line, sizeof line, "%s %d %.6f\n", kv.first.c_str(), kv.second.n, kv.second.ms);
    s += line;
  }
  need(buf && len > s.size(), "otbl_kprof_report: buffer too small");
  memcpy(buf, s.c_str(), s.size() + 1);
  API_END
}

// ---- statistics (stotf.F90:51-291) ----
static const char *const kStatNames[OTBL_NSTAT] = {
    "us", "vs", "ws", "uv", "ua", "va", "wa", "pp", "pm", "vortx", "vorty", "vortz", "vortxa", "vortya",
    "vortza", "uw", "vw", "dispu", "dispv", "dispw", "dispuv", "pup", "pvp", "pdudx", "pdvdy", "pdudy",
    "pdvdx", "pdwdz", "v3", "u2v", "v2u", "w2v", "dudx0", "dudy0", "dvdx0", "dvdy0", "dwdx0", "dwdy0",
    "u3", "w2u"};

const char *otbl_stats_name(int id) { return (id >= 0 && id < OTBL_NSTAT) ? kStatNames[id] : nullptr; }

int otbl_stats_enable(otbl_ctx *ctx, int on) {
  API_BEGIN
  Box &b = box_of(ctx);
  OTBL_HIP(hipStreamSynchronize(b.s));
  drop_graphs(b); // captured samples hold the old buffers
  for (double **p : {&b.st, &b.stcoef, &b.stmean})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  if (b.ens) (void)hipFree(b.ens);
  b.ens = nullptr;
  b.ical = 0;
  b.dostat = false;
  if (on) {
    need(k_stats_shm_bytes(b.nz) <= 160 * 1024 && b.nz <= 4096,
         "otbl_stats_enable: nz above 4096 (the statistics row transforms run in LDS)");
    dalloc(b.st, (size_t)OTBL_NSTAT * b.nx * (b.ny + 1));
    const HostCoef &h = b.hc;
    std::vector<double> c;
    for (const std::vector<double> *v : {&h.fd_dvdy, &h.fd_iy, &h.fd_iyp, &h.fd_dx, &h.fd_ix})
      c.insert(c.end(), v->begin(), v->end());
    dalloc(b.stcoef, c.size());
    h2d(b, b.stcoef, c.data(), c.size() * sizeof(double));
    dalloc(b.ens, (size_t)b.nz * b.ny * b.nx);
    if (b.precision == OTBL_MIXED) dalloc(b.stmean, 5 * (size_t)b.nx * (b.ny + 1));
    k_stats_attrs(b);
  }
  API_END
}

int otbl_stats_schedule(otbl_ctx *ctx, int avcrt, int stest) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(b.st != nullptr, "otbl_stats_schedule: statistics are off (otbl_stats_enable)");
  need(avcrt >= 0, "otbl_stats_schedule: avcrt must be >= 0");
  b.avcrt = avcrt;
  b.stest = stest;
  b.istep = 0;
  API_END
}

int otbl_stats_request(otbl_ctx *ctx) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(b.st != nullptr, "otbl_stats_request: statistics are off (otbl_stats_enable)");
  b.dostat = true;
  API_END
}

int otbl_statsp(otbl_ctx *ctx) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(b.st != nullptr, "otbl_statsp: statistics are off (otbl_stats_enable)");
  k_statsp(b);
  ++b.ical;
  API_END
}

int otbl_stats_get(otbl_ctx *ctx, int id, double *out) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(b.st != nullptr, "otbl_stats_get: statistics are off (otbl_stats_enable)");
  need(id >= 0 && id < OTBL_NSTAT && out, "otbl_stats_get: bad id or null pointer");
  const size_t n = (size_t)b.nx * (b.ny + 1);
  OTBL_HIP(hipStreamSynchronize(b.s));
  d2h(b, out, b.st + (size_t)id * n, n * sizeof(double));
  API_END
}

int otbl_stats_ical(otbl_ctx *ctx, int *ical) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(ical != nullptr, "otbl_stats_ical: null pointer");
  *ical = b.ical;
  API_END
}

int otbl_stats_reset(otbl_ctx *ctx) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(b.st != nullptr, "otbl_stats_reset: statistics are off (otbl_stats_enable)");
  k_stats_zero(b);
  OTBL_HIP(hipStreamSynchronize(b.s));
  b.ical = 0;
  API_END
}

int otbl_stats_get_ens(otbl_ctx *ctx, float *out) {
  API_BEGIN
  Box &b = box_of(ctx);
  need(b.ens != nullptr && out, "otbl_stats_get_ens: statistics are off or null pointer");
  OTBL_HIP(hipStreamSynchronize(b.s));
  d2h(b, out, b.ens, (size_t)b.nz * b.ny * b.nx * sizeof(float));
  API_END
}

} // extern "C"

// ---------------------------------------------------------------------------
// Ensemble of independent boxes: one captured graph per substep whose
// branches are the boxes' substeps (forked from and joined into the
// ensemble's own stream), so the kernels of every box overlap on the GPU.
// ---------------------------------------------------------------------------
struct otbl_ensemble {
  std::vector<Box *> boxes;
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, done = nullptr;
  std::vector<hipEvent_t> join;
  // graph per (substep m, buffer-role bits of the boxes) and the bits of the
  // boxes that take a statistics sample, captured with the boxes' graph
  // generations in gens
  std::map<std::pair<std::pair<int, unsigned long long>, std::pair<int, unsigned long long>>, hipGraphExec_t>
      graphs;
  std::vector<unsigned> gens;
};

namespace {

void ens_drop(otbl_ensemble &e) {
  for (auto &kv : e.graphs) (void)hipGraphExecDestroy(kv.second);
  e.graphs.clear();
}

void ens_free(otbl_ensemble *e) {
  if (!e) return;
  if (e->s) (void)hipStreamSynchronize(e->s);
  ens_drop(*e);
  for (hipEvent_t ev : e->join)
    if (ev) (void)hipEventDestroy(ev);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->done) (void)hipEventDestroy(e->done);
  if (e->s) (void)hipStreamDestroy(e->s);
  delete e;
}

// the ensemble's stream after every box's earlier work
void ens_after_boxes(otbl_ensemble &e) {
  for (size_t q = 0; q < e.boxes.size(); ++q) {
    OTBL_HIP(hipEventRecord(e.join[q], e.boxes[q]->s));
    OTBL_HIP(hipStreamWaitEvent(e.s, e.join[q], 0));
  }
}

// every box's later work after the ensemble's stream
void ens_before_boxes(otbl_ensemble &e) {
  OTBL_HIP(hipEventRecord(e.done, e.s));
  for (Box *b : e.boxes) OTBL_HIP(hipStreamWaitEvent(b->s, e.done, 0));
}

// Substep m of every box (with the end-of-step time update after m = 3).
void ens_substep(otbl_ensemble &e, int m) {
  bool eager = false;
  unsigned long long roles = 0;
  for (size_t q = 0; q < e.boxes.size(); ++q) {
    eager = eager || e.boxes[q]->kprof;
    roles |= (unsigned long long)e.boxes[q]->flip << q;
  }
  unsigned long long smask = 0; // boxes that take a statistics sample
  for (size_t q = 0; q < e.boxes.size(); ++q)
    if (take_stats(*e.boxes[q], m)) smask |= 1ull << q;
  if (eager) { // measurement: kernels one by one on each box's own stream
    for (size_t q = 0; q < e.boxes.size(); ++q) {
      Box *b = e.boxes[q];
      enqueue_substep(*b, m, (smask >> q) & 1ull);
      k_swap_saved(*b);
      if (m == 3) k_end_step(*b);
      if ((smask >> q) & 1ull) stats_taken(*b);
    }
    return;
  }
  for (size_t q = 0; q < e.boxes.size(); ++q)
    if (e.gens[q] != e.boxes[q]->graph_gen) {
      ens_drop(e);
      for (size_t r = 0; r < e.boxes.size(); ++r) e.gens[r] = e.boxes[r]->graph_gen;
      break;
    }
  const std::pair<int, unsigned long long> key{m, roles}, keyst{0, smask};
  hipGraphExec_t &ge = e.graphs[{key, keyst}];
  if (!ge) {
    hipGraph_t g = nullptr;
    size_t done = 0;
    OTBL_HIP(hipStreamBeginCapture(e.s, hipStreamCaptureModeThreadLocal));
    try {
      OTBL_HIP(hipEventRecord(e.fork, e.s));
      for (; done < e.boxes.size(); ++done) {
        Box &b = *e.boxes[done];
        OTBL_HIP(hipStreamWaitEvent(b.s, e.fork, 0));
        enqueue_substep(b, m, (smask >> done) & 1ull);
        if (m == 3) k_end_step(b);
        OTBL_HIP(hipEventRecord(e.join[done], b.s));
        OTBL_HIP(hipStreamWaitEvent(e.s, e.join[done], 0));
      }
    } catch (...) {
      if (hipStreamEndCapture(e.s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
      // capture ran the host side only: undo the role swaps already made
      for (size_t q = 0; q < e.boxes.size(); ++q)
        if (((roles >> q) & 1ull) != (unsigned long long)e.boxes[q]->flip) k_swap_saved(*e.boxes[q]);
      e.graphs.erase({key, keyst});
      throw;
    }
    OTBL_HIP(hipStreamEndCapture(e.s, &g));
    const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    for (size_t q = 0; q < e.boxes.size(); ++q)
      if (((roles >> q) & 1ull) != (unsigned long long)e.boxes[q]->flip) k_swap_saved(*e.boxes[q]);
    if (ie != hipSuccess) {
      e.graphs.erase({key, keyst});
      fail_hip(ie, "hipGraphInstantiate (ensemble substep)");
    }
  }
  OTBL_HIP(hipGraphLaunch(ge, e.s));
  for (size_t q = 0; q < e.boxes.size(); ++q) {
    k_swap_saved(*e.boxes[q]);
    if ((smask >> q) & 1ull) stats_taken(*e.boxes[q]);
  }
}

} // namespace

extern "C" {

int otbl_ensemble_create(otbl_ctx *const *boxes, int nbox, otbl_ensemble **out) {
  otbl_ensemble *e = nullptr;
  try {
  need(out != nullptr && boxes != nullptr && nbox >= 1, "otbl_ensemble_create: bad arguments");
  need(nbox <= 64, "an ensemble holds at most 64 boxes");
  *out = nullptr;
  e = new otbl_ensemble();
  for (int q = 0; q < nbox; ++q) {
    Box &b = box_of(boxes[q]);
    for (Box *o : e->boxes) need(o != &b, "otbl_ensemble_create: a box appears twice");
    e->boxes.push_back(&b);
    e->gens.push_back(b.graph_gen);
  }
  OTBL_HIP(hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking));
  OTBL_HIP(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
  OTBL_HIP(hipEventCreateWithFlags(&e->done, hipEventDisableTiming));
  e->join.assign(nbox, nullptr);
  for (int q = 0; q < nbox; ++q) OTBL_HIP(hipEventCreateWithFlags(&e->join[q], hipEventDisableTiming));
  *out = e;
  return OTBL_OK;
  } catch (const std::invalid_argument &x) {
    ens_free(e);
    g_err = x.what();
    return OTBL_EINVAL;
  } catch (const std::exception &x) {
    ens_free(e);
    g_err = x.what();
    return OTBL_EHIP;
  }
}

int otbl_ensemble_step(otbl_ensemble *e, int nsteps) {
  API_BEGIN
  need(e != nullptr, "null otbl_ensemble");
  need(nsteps >= 0, "nsteps must be >= 0");
  ens_after_boxes(*e);
  try {
    for (int st = 0; st < nsteps; ++st) {
      for (Box *b : e->boxes) begin_step(*b);
      for (int m = 1; m <= 3; ++m) ens_substep(*e, m);
    }
  } catch (...) {
    ens_before_boxes(*e);
    throw;
  }
  ens_before_boxes(*e);
  API_END
}

int otbl_ensemble_sync(otbl_ensemble *e) {
  API_BEGIN
  need(e != nullptr, "null otbl_ensemble");
  OTBL_HIP(hipStreamSynchronize(e->s));
  for (Box *b : e->boxes) {
    OTBL_HIP(hipStreamSynchronize(b->s));
    drain_kprof(*b);
  }
  API_END
}

int otbl_ensemble_destroy(otbl_ensemble *e) {
  ens_free(e);
  return OTBL_OK;
}

} // extern "C"
