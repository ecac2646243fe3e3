// rrtmgsw_api.cpp -- C-ABI entry points of librrtmgsw_hip.so (declared in include/rrtmgsw.h).
//
// Host orchestration only: flag handling and solar-variability scalars of the reference driver
// (src/rrtmg_sw_rad.f90), device buffers and streams, KISS jump-ahead tables, error reporting.
// All per-column, per-g-point work runs in rrtmgsw_kernels.hip; there is no CPU path.
#include <algorithm>
#include <cmath>
#include <tuple>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rrtmgsw.h"
#include "rrtmgsw_internal.h"
#include "rrtmgsw_nc.h"
#include "rrtmgsw_tables.h"

using namespace rrtmgsw;

// Scratch of rrtmgsw_mcica_sw_dev_noalloc: a file-scope device array, one copy per GPU that loads
// the library, held until the process ends. Host code never passes this name to a kernel: its
// device address is looked up once per device (hipGetSymbolAddress, at rrtmgsw_init) and kept in
// the context.
static __device__ __align__(256) char scratch[RRTMGSW_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return fail(RRTMGSW_ERR_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

struct Scratch {
  void *ptr = nullptr;
  size_t bytes = 0;
};

struct Context {
  bool ready = false;
  int device = 0;
  double heatfac = 0.0;
  HostTables host;
  std::string source;             // k-table file the tables came from ("" = compiled-in data)
  DeviceTables dev{};
  double *d_tab = nullptr, *d_exp = nullptr, *d_pthr = nullptr;
  void *d_kisspow = nullptr;      // KissPow: the KISS maps of 2^k draws
  GPoint *d_gp = nullptr;
  std::mutex mu;
  std::map<hipStream_t, Scratch> scratch, status, mtbuf, kissbuf, adjwork, hostbuf;
  std::map<hipStream_t, std::unique_ptr<std::mutex>> calls;   // one call at a time per stream
  // MT19937 jump-ahead tables by segment length (log2): the set bits of x^(k J) mod P for
  // k = 1 .. count, built on first use and extended when a longer stream needs more segments
  struct MtJump {
    int count = 0;
    uint32_t *bits = nullptr, *offs = nullptr;
  };
  std::map<int, MtJump> mtjump;
  std::vector<void *> retired;    // tables replaced by longer ones: a call on another stream may still read them
  std::map<int, char *> noalloc;  // device -> address of `scratch` on that device
  std::mutex noalloc_mu;          // calls of the no-alloc entry share their GPU's array: one at a time
} ctx;

// A call's scratch, status word and generator buffers are its stream's from their acquisition
// to the call's last read-back (every entry ends by synchronising its stream), so calls on one
// stream from different host threads run one at a time; calls on different streams overlap.
std::unique_lock<std::mutex> stream_call(hipStream_t s) {
  std::mutex *m;
  {
    std::lock_guard<std::mutex> lk(ctx.mu);
    auto &p = ctx.calls[s];
    if (!p) p.reset(new std::mutex);
    m = p.get();
  }
  return std::unique_lock<std::mutex>(*m);
}

// Grow-only per-stream device buffer.
int stream_buffer(std::map<hipStream_t, Scratch> &m, hipStream_t s, size_t bytes, void **out) {
  Scratch &b = m[s];
  if (b.bytes < bytes) {
    if (b.ptr) HIPCHK(hipFree(b.ptr));
    b.ptr = nullptr;
    HIPCHK(hipMalloc(&b.ptr, bytes));
    b.bytes = bytes;
  }
  *out = b.ptr;
  return 0;
}

// ---------------- host-array transfers ----------------
// The host-array entry points take the caller's ordinary (pageable) arrays. Transfers pass
// through two pinned staging buffers allocated once (hipHostMalloc): while the DMA engine moves
// one chunk, host threads copy the next chunk (host to device) or the previous one (device to
// host) between the caller's array and the other buffer.
constexpr size_t STAGE_BYTES = (size_t)32 << 20;

// A fixed set of host threads for large copies, started on first use, never stopped (they wait
// on a condition variable when idle).
struct CopyPool {
  std::mutex m;
  std::condition_variable go, done;
  int nth = 0, pending = 0;
  uint64_t gen = 0;
  char *dst = nullptr;
  const char *src = nullptr;
  size_t n = 0, part = 0;
  void start() {
    const unsigned hw = std::thread::hardware_concurrency();
    nth = (int)std::min<unsigned>(15, hw > 1 ? hw - 1 : 0);
    for (int t = 1; t <= nth; t++) std::thread([this, t] { worker(t); }).detach();
  }
  void slice(int t, char *d, const char *s2, size_t total, size_t p) {
    const size_t o = (size_t)t * p;
    if (o < total) memcpy(d + o, s2 + o, std::min(p, total - o));
  }
  void worker(int t) {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
      go.wait(lk, [&] { return gen != seen; });
      seen = gen;
      char *d = dst;
      const char *s2 = src;
      const size_t total = n, p = part;
      lk.unlock();
      slice(t, d, s2, total, p);
      lk.lock();
      if (--pending == 0) done.notify_one();
    }
  }
  void copy(void *d, const void *s2, size_t bytes) {
    if (nth == 0 || bytes < ((size_t)4 << 20)) {
      memcpy(d, s2, bytes);
      return;
    }
    std::unique_lock<std::mutex> lk(m);
    dst = (char *)d, src = (const char *)s2, n = bytes;
    part = ((bytes / (nth + 1)) + 4095) & ~(size_t)4095;
    pending = nth;
    gen++;
    go.notify_all();
    lk.unlock();
    slice(0, dst, src, bytes, part);
    lk.lock();
    done.wait(lk, [&] { return pending == 0; });
  }
};

struct Stager {
  std::mutex mu;   // one host-array transfer at a time
  char *pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2];
  bool live[2] = {false, false};
  bool ready = false;
  CopyPool *pool = nullptr;   // leaked on purpose: its threads outlive any static destructor
} stager;

int stager_init() {
  if (stager.ready) return 0;
  for (int b = 0; b < 2; b++) {
    HIPCHK(hipHostMalloc((void **)&stager.pin[b], STAGE_BYTES, hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&stager.ev[b], hipEventDisableTiming));
  }
  stager.pool = new CopyPool();
  stager.pool->start();
  stager.ready = true;
  return 0;
}
int stage_wait(int b) {
  if (stager.live[b]) {
    HIPCHK(hipEventSynchronize(stager.ev[b]));
    stager.live[b] = false;
  }
  return 0;
}
// host -> device, queued on s ahead of the work queued after it
int staged_h2d(void *dev, const void *host, size_t bytes, hipStream_t s) {
  size_t i = 0;
  for (size_t off = 0; off < bytes; off += STAGE_BYTES, i++) {
    const int b = (int)(i & 1);
    const size_t n = std::min(STAGE_BYTES, bytes - off);
    if (int rc = stage_wait(b)) return rc;
    stager.pool->copy(stager.pin[b], (const char *)host + off, n);
    HIPCHK(hipMemcpyAsync((char *)dev + off, stager.pin[b], n, hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(stager.ev[b], s));
    stager.live[b] = true;
  }
  return 0;
}
// device -> host after the work queued on s before it; returns with the data in host memory
int staged_d2h(void *host, const void *dev, size_t bytes, hipStream_t s) {
  size_t i = 0, prev_off = 0, prev_n = 0;
  int prev_b = -1;
  for (size_t off = 0; off < bytes; off += STAGE_BYTES, i++) {
    const int b = (int)(i & 1);
    const size_t n = std::min(STAGE_BYTES, bytes - off);
    if (int rc = stage_wait(b)) return rc;
    HIPCHK(hipMemcpyAsync(stager.pin[b], (const char *)dev + off, n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(stager.ev[b], s));
    stager.live[b] = true;
    if (prev_b >= 0) {
      if (int rc = stage_wait(prev_b)) return rc;
      stager.pool->copy((char *)host + prev_off, stager.pin[prev_b], prev_n);
    }
    prev_b = b, prev_off = off, prev_n = n;
  }
  if (prev_b >= 0) {
    if (int rc = stage_wait(prev_b)) return rc;
    stager.pool->copy((char *)host + prev_off, stager.pin[prev_b], prev_n);
  }
  return 0;
}

// The host-array form of an entry point: the inputs that are present staged into one persistent
// device buffer, the device entry run (run(inputs, outputs), on the null stream, which it
// synchronizes), the outputs staged back. An input with n > 0 and no array is an error.
struct HostIn {
  const double *h;
  size_t n;
};
struct HostOut {
  double *h;
  size_t n;
};
template <typename F>
int host_call(const HostIn *in, int nin, const HostOut *out, int nout, F run) {
  std::lock_guard<std::mutex> hl(stager.mu);
  if (int rc = stager_init()) return rc;
  size_t total = 0;
  for (int i = 0; i < nin; i++) total += in[i].n;
  for (int i = 0; i < nout; i++) total += out[i].n;
  void *dv = nullptr;
  {
    std::lock_guard<std::mutex> lk(ctx.mu);
    if (int rc = stream_buffer(ctx.hostbuf, nullptr, std::max<size_t>(total, 1) * sizeof(double), &dv)) return rc;
  }
  double *d = (double *)dv;
  std::vector<const double *> din(nin);
  std::vector<double *> dout(nout);
  size_t off = 0;
  for (int i = 0; i < nin; i++) {
    din[i] = in[i].n ? d + off : nullptr;
    if (in[i].n) {
      if (!in[i].h) return fail(RRTMGSW_ERR_ARGUMENT, "a required input array is NULL");
      if (int rc = staged_h2d(d + off, in[i].h, in[i].n * sizeof(double), nullptr)) return rc;
    }
    off += in[i].n;
  }
  for (int i = 0; i < nout; i++) dout[i] = d + off, off += out[i].n;
  int rc = run(din.data(), dout.data());
  for (int i = 0; rc == 0 && i < nout; i++)
    if (out[i].n) rc = staged_d2h(out[i].h, dout[i], out[i].n * sizeof(double), nullptr);
  return rc;
}

// ---------------- KISS jump-ahead (host) ----------------
// Sub-column g of a column starts its draws at position W + g*D of the column's stream
// (W = permuteseed warm-up draws, D = draws per sub-column). Each KISS component has a
// closed-form jump: the LCG s1 is affine, the xorshift s2 is linear over GF(2), and the
// multiply-with-carry s3/s4 are multiplications by 2^-16 = a modulo a*2^16-1.
struct Affine {
  uint32_t a, c;
};
Affine affine_compose(Affine f, Affine g) {  // f(g(x))
  return {f.a * g.a, f.a * g.c + f.c};
}
Affine affine_pow(Affine f, uint64_t n) {
  Affine r = {1u, 0u};
  while (n) {
    if (n & 1) r = affine_compose(f, r);
    f = affine_compose(f, f);
    n >>= 1;
  }
  return r;
}
struct Gf2 {
  uint32_t col[32];
};
uint32_t gf2_apply(const Gf2 &m, uint32_t x) {
  uint32_t r = 0;
  for (int i = 0; i < 32; i++)
    if ((x >> i) & 1u) r ^= m.col[i];
  return r;
}
Gf2 gf2_compose(const Gf2 &a, const Gf2 &b) {  // a(b(x))
  Gf2 r;
  for (int i = 0; i < 32; i++) r.col[i] = gf2_apply(a, b.col[i]);
  return r;
}
Gf2 gf2_pow(Gf2 m, uint64_t n) {
  Gf2 r;
  for (int i = 0; i < 32; i++) r.col[i] = 1u << i;
  while (n) {
    if (n & 1) r = gf2_compose(m, r);
    m = gf2_compose(m, m);
    n >>= 1;
  }
  return r;
}
uint32_t powmod(uint64_t b, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  b %= m;
  while (e) {
    if (e & 1) r = (r * b) % m;
    b = (b * b) % m;
    e >>= 1;
  }
  return (uint32_t)r;
}

// The one-draw maps of the four kissvec components (src/mcica_subcol_gen_sw.f90:721-753).
Gf2 kiss_xorshift() {
  Gf2 xs;
  for (int i = 0; i < 32; i++) {
    uint32_t k = 1u << i;
    k ^= k << 13;
    k ^= k >> 17;
    k ^= k << 5;
    xs.col[i] = k;
  }
  return xs;
}
const Affine KISS_LCG = {69069u, 1327217885u};
constexpr uint64_t KISS_M3 = 1179647999ull, KISS_M4 = 2025259007ull;   // 18000 2^16 - 1, 30903 2^16 - 1

// The state map of n kissvec draws: an affine map for the congruential part, a GF(2) matrix
// power for the xorshift, modular powers for the two multiply-with-carry parts.
KissJump kiss_jump_of(uint64_t n) {
  KissJump J;
  Affine a = affine_pow(KISS_LCG, n);
  Gf2 m = gf2_pow(kiss_xorshift(), n);
  J.a1 = a.a, J.c1 = a.c;
  memcpy(J.m2, m.col, sizeof m.col);
  J.p3 = powmod(18000, n, KISS_M3);
  J.p4 = powmod(30903, n, KISS_M4);
  return J;
}

// The maps of 2^k draws, k = 0..63, built once at initialisation: the jump of n draws is the
// composition of those of n's set bits (all powers of one map commute).
struct KissPow {
  uint32_t m2[64][32];   // xorshift matrix columns
  uint32_t a1[64], c1[64], p3[64], p4[64];
};
KissPow kiss_pow_table() {
  KissPow P;
  Gf2 m = kiss_xorshift();
  Affine a = KISS_LCG;
  uint64_t p3 = 18000, p4 = 30903;
  for (int k = 0; k < 64; k++) {
    memcpy(P.m2[k], m.col, sizeof m.col);
    P.a1[k] = a.a, P.c1[k] = a.c, P.p3[k] = (uint32_t)p3, P.p4[k] = (uint32_t)p4;
    m = gf2_compose(m, m), a = affine_compose(a, a);
    p3 = p3 * p3 % KISS_M3, p4 = p4 * p4 % KISS_M4;
  }
  return P;
}

// The jump records of the 112 sub-columns, n = w + g per_sub draws into the column's stream:
// 32 lanes per g-point, lane i forming column i of the xorshift matrix by applying the maps of
// n's set bits (k < kmax, the bit length of the largest n) to the identity's column i; lane 0
// also forms the affine and multiply-with-carry parts.
__global__ void __launch_bounds__(64) kiss_jumps_kernel(const KissPow *P, uint64_t w, uint64_t per_sub,
                                                        int kmax, KissJump *out) {
  const int lane = threadIdx.x & 31, g = blockIdx.x * 2 + (threadIdx.x >> 5);
  if (g >= NGPT) return;
  const uint64_t n = w + (uint64_t)g * per_sub;
  uint32_t col = 1u << lane, a = 1u, c = 0u;
  uint64_t q3 = 1, q4 = 1;
  for (int k = 0; k < kmax; k++) {
    if (!((n >> k) & 1u)) continue;
    const uint32_t *m = P->m2[k];
    uint32_t r = 0;
    for (int j = 0; j < 32; j++) r ^= m[j] & (0u - ((col >> j) & 1u));
    col = r;
    c = P->a1[k] * c + P->c1[k], a = P->a1[k] * a;
    q3 = q3 * P->p3[k] % KISS_M3, q4 = q4 * P->p4[k] % KISS_M4;
  }
  out[g].m2[lane] = col;
  if (lane == 0) out[g].a1 = a, out[g].c1 = c, out[g].p3 = (uint32_t)q3, out[g].p4 = (uint32_t)q4;
}

// The per-sub-column KISS jumps for (permuteseed, draws per sub-column), formed on the GPU into
// `buf` ahead of the work queued after it on s: no host copy, no wait, no cache.
int kiss_jumps(int permuteseed, int per_sub, hipStream_t s, KissJump *buf) {
  const uint64_t w = permuteseed > 0 ? (uint64_t)permuteseed : 0;
  const uint64_t nmax = w + (uint64_t)(NGPT - 1) * (uint64_t)per_sub;
  const int kmax = nmax ? 64 - __builtin_clzll(nmax) : 0;
  hipLaunchKernelGGL(kiss_jumps_kernel, dim3((NGPT + 1) / 2), dim3(64), 0, s,
                     (const KissPow *)ctx.d_kisspow, w, (uint64_t)per_sub, kmax, buf);
  HIPCHK(hipGetLastError());
  return 0;
}
int kiss_jumps(int permuteseed, int per_sub, hipStream_t s, const uint32_t **out) {
  void *buf = nullptr;
  {
    std::lock_guard<std::mutex> lk(ctx.mu);
    if (int rc = stream_buffer(ctx.kissbuf, s, sizeof(KissJump) * NGPT, &buf)) return rc;
  }
  *out = (const uint32_t *)buf;
  return kiss_jumps(permuteseed, per_sub, s, (KissJump *)buf);
}

// ---------------- solar variability and flux adjustment (host) ----------------
// inatm_sw, src/rrtmg_sw_rad.f90:1226-1512; earth_sun :844-868.
const double MGAVGCYC[134] = {
    0.150737, 0.150746, 0.150733, 0.150718, 0.150725, 0.150762, 0.150828, 0.150918, 0.151017, 0.151113,
    0.151201, 0.151292, 0.151403, 0.151557, 0.151766, 0.152023, 0.152322, 0.152646, 0.152969, 0.153277,
    0.153579, 0.153899, 0.154252, 0.154651, 0.155104, 0.155608, 0.156144, 0.156681, 0.157178, 0.157605,
    0.157971, 0.158320, 0.158702, 0.159133, 0.159583, 0.160018, 0.160408, 0.160725, 0.160960, 0.161131,
    0.161280, 0.161454, 0.161701, 0.162034, 0.162411, 0.162801, 0.163186, 0.163545, 0.163844, 0.164029,
    0.164054, 0.163910, 0.163621, 0.163239, 0.162842, 0.162525, 0.162344, 0.162275, 0.162288, 0.162369,
    0.162500, 0.162671, 0.162878, 0.163091, 0.163251, 0.163320, 0.163287, 0.163153, 0.162927, 0.162630,
    0.162328, 0.162083, 0.161906, 0.161766, 0.161622, 0.161458, 0.161266, 0.161014, 0.160666, 0.160213,
    0.159690, 0.159190, 0.158831, 0.158664, 0.158634, 0.158605, 0.158460, 0.158152, 0.157691, 0.157152,
    0.156631, 0.156180, 0.155827, 0.155575, 0.155406, 0.155280, 0.155145, 0.154972, 0.154762, 0.154554,
    0.154388, 0.154267, 0.154152, 0.154002, 0.153800, 0.153567, 0.153348, 0.153175, 0.153044, 0.152923,
    0.152793, 0.152652, 0.152510, 0.152384, 0.152282, 0.152194, 0.152099, 0.151980, 0.151844, 0.151706,
    0.151585, 0.151496, 0.151437, 0.151390, 0.151347, 0.151295, 0.151220, 0.151115, 0.150993, 0.150883,
    0.150802, 0.150752, 0.150729, 0.150737};
const double SBAVGCYC[134] = {
    50.3550, 44.1322, 52.0179, 59.2231, 66.3702, 71.7545, 76.8671, 83.4723, 91.1574, 98.4915,
    105.3173, 115.1791, 130.9432, 155.0483, 186.5379, 221.5456, 256.9212, 291.5276, 325.2953, 356.4789,
    387.2470, 422.8557, 466.1698, 521.5139, 593.2833, 676.6234, 763.6930, 849.1200, 928.4259, 994.9705,
    1044.2605, 1087.5703, 1145.0623, 1224.3491, 1320.6497, 1413.0979, 1472.1591, 1485.7531, 1464.1610, 1439.1617,
    1446.2449, 1496.4323, 1577.8394, 1669.5933, 1753.0408, 1821.9296, 1873.2789, 1906.5240, 1920.4482, 1904.6881,
    1861.8397, 1802.7661, 1734.0215, 1665.0562, 1608.8999, 1584.8208, 1594.0162, 1616.1486, 1646.6031, 1687.1962,
    1736.4778, 1787.2419, 1824.9084, 1835.5236, 1810.2161, 1768.6124, 1745.1085, 1748.7762, 1756.1239, 1738.9929,
    1700.0656, 1658.2209, 1629.2925, 1620.9709, 1622.5157, 1623.4703, 1612.3083, 1577.3031, 1516.7953, 1430.0403,
    1331.5112, 1255.5171, 1226.7653, 1241.4419, 1264.6549, 1255.5559, 1203.0286, 1120.2747, 1025.5101, 935.4602,
    855.0434, 781.0189, 718.0328, 678.5850, 670.4219, 684.1906, 697.0376, 694.8083, 674.1456, 638.8199,
    602.3454, 577.6292, 565.6213, 553.7846, 531.7452, 503.9732, 476.9708, 452.4296, 426.2826, 394.6636,
    360.1086, 324.9731, 297.2957, 286.1536, 287.4195, 288.9029, 282.7594, 267.7211, 246.6594, 224.7318,
    209.2318, 204.5217, 204.1653, 200.0440, 191.0689, 175.7699, 153.9869, 128.4389, 103.8445, 85.6083,
    73.6264, 64.4393, 56.5779, 50.3550};

void solar_scalars(double adjes, int dyofyr, double scon, int isolvar, const double *bndsolvar,
                   const double *indsolvar, const double *solcycfrac, SwScalars &o) {
  const double svar_f_avg = 0.1567652, svar_s_avg = 902.71260;
  const double Iint = 1360.37, Fint = 0.996047, Sint = -0.511590;
  const double Foffset = 0.14959542, Soffset = 0.00066696;
  const double cyc_min = 0.0189, cyc_max = 0.3750, d_min2max = 0.3561, d_max2min = 0.6439;
  const double rrsw_scon = 1.36822e+03;
  const int nsolfrac = 134;
  double solvar[NBND];
  for (int b = 0; b < NBND; b++) {
    solvar[b] = 1.0;
    o.svar_bnd[0][b] = o.svar_bnd[1][b] = o.svar_bnd[2][b] = 1.0;
  }
  o.svar_f = o.svar_s = o.svar_i = 1.0;
  double scl[2] = {1.0, 1.0}, ndx[2] = {svar_f_avg, svar_s_avg}, cycfr = 0.0;
  if (isolvar == 1 && indsolvar && solcycfrac) {  // :1250-1276
    cycfr = *solcycfrac;
    if (indsolvar[0] != 1.0 || indsolvar[1] != 1.0) {
      double f = *solcycfrac, w;
      if (f >= 0.0 && f < cyc_min) {
        w = (f + 1.0 - cyc_max) / d_max2min;
        scl[0] = indsolvar[0] + w * (1.0 - indsolvar[0]);
        scl[1] = indsolvar[1] + w * (1.0 - indsolvar[1]);
      }
      if (f >= cyc_min && f <= cyc_max) {
        w = (f - cyc_min) / d_min2max;
        scl[0] = 1.0 + w * (indsolvar[0] - 1.0);
        scl[1] = 1.0 + w * (indsolvar[1] - 1.0);
      }
      if (f > cyc_max && f <= 1.0) {
        w = (f - cyc_max) / d_max2min;
        scl[0] = indsolvar[0] + w * (1.0 - indsolvar[0]);
        scl[1] = indsolvar[1] + w * (1.0 - indsolvar[1]);
      }
    }
  }
  if (isolvar == 2 && indsolvar) ndx[0] = indsolvar[0], ndx[1] = indsolvar[1];  // :1279-1287
  double adjflx = dyofyr > 0 ? 0.0 : adjes;                                      // :1291-1294
  if (dyofyr > 0) {
    double pi = 2.0 * asin(1.0), gamma = 2.0 * pi * (dyofyr - 1) / 365.0;
    adjflx = 1.000110 + .034221 * cos(gamma) + .001289 * sin(gamma) + .000719 * cos(2.0 * gamma) +
             .000077 * sin(2.0 * gamma);
  }
  double f0 = 0.0, s0 = 0.0;
  if (isolvar == 1) {  // mean-cycle interpolation, :1323-1351 / :1411-1439
    if (cycfr <= 0.0) {
      f0 = MGAVGCYC[0], s0 = SBAVGCYC[0];
    } else if (cycfr >= 1.0) {
      f0 = MGAVGCYC[nsolfrac - 1], s0 = SBAVGCYC[nsolfrac - 1];
    } else {
      double len = 1.0 / (nsolfrac - 2), hf = 0.5 * len, lo = 0.0, hi = hf;
      int sfid = 1;
      if (cycfr > hf && cycfr < 1.0 - hf) {
        sfid = (int)floor((cycfr - hf) * (nsolfrac - 2)) + 2;
        lo = (sfid - 2) * len + hf;
        hi = lo + len;
      }
      if (cycfr >= 1.0 - hf) sfid = nsolfrac - 1, lo = 1.0 - hf, hi = 1.0;
      double fr = (cycfr - lo) / (hi - lo);
      f0 = MGAVGCYC[sfid - 1] + fr * (MGAVGCYC[sfid] - MGAVGCYC[sfid - 1]);
      s0 = SBAVGCYC[sfid - 1] + fr * (SBAVGCYC[sfid] - SBAVGCYC[sfid - 1]);
    }
  }
  if (scon == 0.0) {  // :1301-1380
    if (isolvar == -1 || isolvar == 3)
      for (int b = 0; b < NBND; b++) solvar[b] = bndsolvar ? bndsolvar[b] : 1.0;
    if (isolvar == 1) {
      o.svar_f = scl[0] * (f0 - Foffset) / (svar_f_avg - Foffset);
      o.svar_s = scl[1] * (s0 - Soffset) / (svar_s_avg - Soffset);
    }
    if (isolvar == 2) {
      o.svar_f = (ndx[0] - Foffset) / (svar_f_avg - Foffset);
      o.svar_s = (ndx[1] - Soffset) / (svar_s_avg - Soffset);
    }
  } else if (scon > 0.0) {  // :1384-1479
    double cprim = Fint + Sint + Iint;
    if (isolvar == -1)
      for (int b = 0; b < NBND; b++) solvar[b] = bndsolvar ? bndsolvar[b] * scon / rrsw_scon : scon / rrsw_scon;
    if (isolvar == 0) o.svar_f = o.svar_s = o.svar_i = scon / cprim;
    if (isolvar == 1) {
      o.svar_i = (scon - (scl[0] * Fint + scl[1] * Sint)) / Iint;
      o.svar_f = scl[0] * (f0 - Foffset) / (svar_f_avg - Foffset);
      o.svar_s = scl[1] * (s0 - Soffset) / (svar_s_avg - Soffset);
    }
    if (isolvar == 3)
      for (int b = 0; b < NBND; b++) solvar[b] = bndsolvar ? bndsolvar[b] * scon / cprim : scon / cprim;
  }
  if (isolvar == 3)
    for (int b = 0; b < NBND; b++) o.svar_bnd[0][b] = o.svar_bnd[1][b] = o.svar_bnd[2][b] = solvar[b];
  for (int b = 0; b < NBND; b++) o.adjflux[b] = isolvar < 0 ? adjflx * solvar[b] : adjflx;
  o.isolvar = isolvar;
}

__global__ void zero_int(int *p) { *p = 0; }

int finish(hipStream_t s, int *d_status) {
  int st = 0;
  HIPCHK(hipMemcpyAsync(&st, d_status, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  // the reference's STOP messages of the cloud optics, by CloudStop
  static const char *const cloud_msg[CS_COUNT] = {
      nullptr, "ICE RADIUS OUT OF BOUNDS", "ICE GENERALIZED EFFECTIVE SIZE OUT OF BOUNDS",
      "LIQUID EFFECTIVE RADIUS OUT OF BOUNDS", "FDELTA LESS THAN 0.0", "FDELTA GT THAN 1.0",
      "ICE EXTINCTION LESS THAN 0.0", "ICE SSA GRTR THAN 1.0", "ICE SSA LESS THAN 0.0",
      "ICE ASYM GRTR THAN 1.0", "ICE ASYM LESS THAN 0.0", "LIQUID EXTINCTION LESS THAN 0.0",
      "LIQUID SSA GRTR THAN 1.0", "LIQUID SSA LESS THAN 0.0", "LIQUID ASYM GRTR THAN 1.0",
      "LIQUID ASYM LESS THAN 0.0"};
  const int which = st >> 8;
  st &= 0xff;
  switch (st) {
    case 0: return 0;
    case RRTMGSW_ERR_INFLAG: return fail(st, "INFLAG = 1 OPTION NOT AVAILABLE WITH MCICA");
    case RRTMGSW_ERR_RADIUS:
    case RRTMGSW_ERR_OPTICS:
      if (which > 0 && which < CS_COUNT) return fail(st, cloud_msg[which]);
      return fail(st, st == RRTMGSW_ERR_RADIUS ? "CLOUD PARTICLE RADIUS OUT OF BOUNDS"
                                               : "CLOUD OPTICAL PROPERTY OUT OF PHYSICAL LIMITS");
    case RRTMGSW_ERR_PARTIAL_CLOUD: return fail(st, "PARTIAL CLOUD NOT ALLOWED");
    case RRTMGSW_ERR_KISS_SEED:
      return fail(st, "MCICA_SUBCOL: KISSVEC SEED GENERATOR REQUIRES PMID FROM BOTTOM FOUR LAYERS.");
    default: return fail(st, "kernel error");
  }
}

// Columns per launch so that the per-column elimination coefficients stay within ~4 GiB.
// RRTMGSW_MAX_LAUNCH_COLUMNS (environment, read per call) lowers the columns per launch further,
// so that the multi-launch path can be exercised at small sizes.
size_t launch_column_limit(size_t n) {
  if (const char *e = std::getenv("RRTMGSW_MAX_LAUNCH_COLUMNS")) {
    long v = std::atol(e);
    if (v > 0) n = std::min<size_t>(n, (size_t)v);
  }
  return n;
}

int chunk_columns(int ncol, int nlay) {
  size_t per = scr_alloc_doubles(nlay) * sizeof(double);
  size_t cap = (size_t)4 << 30;
  size_t n = cap / per;
  if (n < 64) n = 64;
  return (int)std::min<size_t>((size_t)ncol, launch_column_limit(n));
}

struct SwCall {
  int ncol, nlay, icld, iaer;
  const double *play, *plev, *tlay, *h2ovmr, *o3vmr, *co2vmr, *ch4vmr, *o2vmr;
  const double *asdir, *asdif, *aldir, *aldif, *coszen;
  const double *cldfmcl, *taucmcl, *ssacmcl, *asmcmcl, *fsfcmcl, *ciwpmcl, *clwpmcl, *reicmcl, *relqmcl;
  const double *tauaer, *ssaaer, *asmaer, *ecaer;
  double *swuflx, *swdflx, *swhr, *swuflxc, *swdflxc, *swhrc;
};


// The status word cleared (by a kernel), then the forward kernel of the given form over column
// chunks of at most `chunk` columns, all on stream s, into a.scratch / a.status.
hipError_t launch_rt_chunks(RtArgs a, const DeviceTables &tables, hipStream_t s, int precision, int chunk,
                            bool spectral = false) {
  hipLaunchKernelGGL(zero_int, dim3(1), dim3(1), 0, s, a.status);
  hipError_t e = hipGetLastError();
  for (int c0 = 0; c0 < a.ncol && e == hipSuccess; c0 += chunk) {
    a.col0 = c0;
    a.ncol_chunk = std::min(chunk, a.ncol - c0);
    e = spectral                        ? launch_rt_spectral(a, tables, s)
        : precision == RRTMGSW_MIXED     ? launch_rt_mixed(a, tables, s)
        : precision == RRTMGSW_FP64_EXP ? launch_rt_exact(a, tables, s)
                                        : launch_rt_fp64(a, tables, s);
  }
  return e;
}

// spectral: the column program's per-band outputs (a.bandu/bandd/banddir); the per-g level terms
// of a chunk go to a buffer allocated for the call.
int run_rt(const RtArgs &base, const DeviceTables &tables, hipStream_t s, int precision = RRTMGSW_FP64,
           bool spectral = false) {
  int chunk = chunk_columns(base.ncol, base.nlay);
  void *scr = nullptr, *st = nullptr;
  {
    std::lock_guard<std::mutex> lk(ctx.mu);
    if (int rc = stream_buffer(ctx.scratch, s, (size_t)chunk * scr_alloc_doubles(base.nlay) * sizeof(double), &scr)) return rc;
    if (int rc = stream_buffer(ctx.status, s, 256, &st)) return rc;
  }
  RtArgs a = base;
  if (spectral) {
    size_t per = (size_t)(base.nlay + 1) * 4 * NGPT;
    chunk = (int)std::min<size_t>((size_t)chunk, std::max<size_t>(1, ((size_t)1 << 30) / (per * sizeof(double))));
    HIPCHK(hipMalloc(&a.gspec, (size_t)chunk * per * sizeof(double)));
  }
  a.scratch = (double *)scr;
  a.status = (int *)st;
  hipError_t e = launch_rt_chunks(a, tables, s, precision, chunk, spectral);
  int rc = e != hipSuccess ? fail(RRTMGSW_ERR_HIP, hipGetErrorString(e)) : finish(s, (int *)st);
  if (spectral) {
    hipStreamSynchronize(s);
    hipFree(a.gspec);
  }
  return rc;
}

int check_init() {
  if (!ctx.ready) return fail(RRTMGSW_ERR_NOT_INITIALIZED, "rrtmgsw_init has not been called");
  return 0;
}

}  // namespace

extern "C" {

const char *rrtmgsw_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace {

// rrtmgsw_init / rrtmgsw_init_nc. `source`: the k-table file, or "" for the compiled-in tables.
int init_tables(double cpdair, const std::string &source) {
  std::lock_guard<std::mutex> lk(ctx.mu);
  if (!(cpdair > 0.0)) return fail(RRTMGSW_ERR_ARGUMENT, "cpdair must be positive");
  if (ctx.ready && ctx.source != source)
    return fail(RRTMGSW_ERR_ARGUMENT,
                "the tables are already initialised from " +
                    (ctx.source.empty() ? std::string("the compiled-in data") : ctx.source) +
                    ": call rrtmgsw_finalize first");
  ctx.heatfac = 9.8066 * 8.6400e4 / (cpdair * 1.e2);  // src/rrtmg_sw_init.f90:258
  if (ctx.ready) return 0;
  std::string err;
  if (!build_tables(ctx.host, err, source.empty() ? nullptr : source.c_str()))
    return fail(RRTMGSW_ERR_ARGUMENT, err);
  HIPCHK(hipGetDevice(&ctx.device));
  const HostTables &H = ctx.host;
  HIPCHK(hipMalloc(&ctx.d_tab, H.tab.size() * sizeof(double)));
  HIPCHK(hipMemcpy(ctx.d_tab, H.tab.data(), H.tab.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&ctx.d_exp, H.exp_tbl.size() * sizeof(double)));
  HIPCHK(hipMemcpy(ctx.d_exp, H.exp_tbl.data(), H.exp_tbl.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&ctx.d_pthr, H.pthr.size() * sizeof(double)));
  HIPCHK(hipMemcpy(ctx.d_pthr, H.pthr.data(), H.pthr.size() * sizeof(double), hipMemcpyHostToDevice));
  {
    const KissPow P = kiss_pow_table();
    HIPCHK(hipMalloc(&ctx.d_kisspow, sizeof P));
    HIPCHK(hipMemcpy(ctx.d_kisspow, &P, sizeof P, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMalloc(&ctx.d_gp, H.gp.size() * sizeof(GPoint)));
  HIPCHK(hipMemcpy(ctx.d_gp, H.gp.data(), H.gp.size() * sizeof(GPoint), hipMemcpyHostToDevice));
  DeviceTables &D = ctx.dev;
  D.tab = ctx.d_tab, D.exp_tbl = ctx.d_exp, D.gp = ctx.d_gp, D.pthr = ctx.d_pthr;
  D.extliq1 = H.extliq1, D.ssaliq1 = H.ssaliq1, D.asyliq1 = H.asyliq1;
  D.extice2 = H.extice2, D.ssaice2 = H.ssaice2, D.asyice2 = H.asyice2;
  D.extice3 = H.extice3, D.ssaice3 = H.ssaice3, D.asyice3 = H.asyice3, D.fdlice3 = H.fdlice3;
  D.abari = H.abari, D.bbari = H.bbari, D.cbari = H.cbari, D.dbari = H.dbari;
  D.ebari = H.ebari, D.fbari = H.fbari;
  D.rsrtaua = H.rsrtaua, D.rsrpiza = H.rsrpiza, D.rsrasya = H.rsrasya;
  D.kiss_jump = nullptr;
  {
    char *p = nullptr;
    HIPCHK(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
    ctx.noalloc[ctx.device] = p;
  }
  ctx.source = source;
  ctx.ready = true;
  return 0;
}

}  // namespace

extern "C" {

int rrtmgsw_init(double cpdair) { return init_tables(cpdair, std::string()); }

int rrtmgsw_init_nc(double cpdair, const char *path) {
  return init_tables(cpdair, path && *path ? std::string(path) : std::string("rrtmg_sw.nc"));
}

int rrtmgsw_host_tables(const char *ktable_file, double *out, size_t cap, size_t *n) {
  HostTables H;
  std::string err;
  if (!build_tables(H, err, ktable_file)) return fail(RRTMGSW_ERR_ARGUMENT, err);
  std::vector<double> v(H.tab);
  for (const GPoint &g : H.gp)
    v.insert(v.end(), {g.rayl_lo, g.rayl_hi, g.lo_minor_coef, g.hi_minor_coef});
  if (n) *n = v.size();
  if (out) std::copy(v.begin(), v.begin() + std::min(cap, v.size()), out);
  return 0;
}

int rrtmgsw_nc_get_vara_double(const char *path, const char *name, int ndims, const size_t *start,
                               const size_t *count, double *out) {
  if (!path || !name || ndims < 0 || (ndims && (!start || !count)))
    return fail(RRTMGSW_ERR_ARGUMENT, "rrtmgsw_nc_get_vara_double: null argument");
  NcFile nc;
  std::string err;
  if (!nc.open(path, err)) return fail(RRTMGSW_ERR_ARGUMENT, err);
  const NcFile::Var *v = nc.var(name);
  if (!v) return fail(RRTMGSW_ERR_ARGUMENT, std::string(path) + ": no variable " + name);
  if ((int)v->shape.size() != ndims)
    return fail(RRTMGSW_ERR_ARGUMENT, std::string(path) + ": variable " + name + " has " +
                                          std::to_string(v->shape.size()) + " dimensions");
  std::vector<uint64_t> s(start, start + ndims), c(count, count + ndims);
  if (!nc.get(*v, s.data(), c.data(), out, err)) return fail(RRTMGSW_ERR_ARGUMENT, err);
  return 0;
}

int rrtmgsw_finalize(void) {
  std::lock_guard<std::mutex> lk(ctx.mu);
  if (!ctx.ready) return 0;
  hipFree(ctx.d_tab);
  hipFree(ctx.d_exp);
  hipFree(ctx.d_gp);
  hipFree(ctx.d_pthr);
  hipFree(ctx.d_kisspow);
  for (auto *m : {&ctx.scratch, &ctx.status, &ctx.mtbuf, &ctx.kissbuf, &ctx.adjwork, &ctx.hostbuf}) {
    for (auto &kv : *m) hipFree(kv.second.ptr);
    m->clear();
  }
  for (auto &kv : ctx.mtjump) hipFree(kv.second.bits), hipFree(kv.second.offs);
  ctx.mtjump.clear();
  for (void *p : ctx.retired) hipFree(p);
  ctx.retired.clear();
  ctx.ready = false;
  return 0;
}

int rrtmgsw_sw_dev(int ncol, int nlay, int *icld, int *iaer, const double *play, const double *plev,
                   const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
                   const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                   const double *n2ovmr, const double *o2vmr, const double *asdir,
                   const double *asdif, const double *aldir, const double *aldif,
                   const double *coszen, double adjes, int dyofyr, double scon, int isolvar,
                   int inflgsw, int iceflgsw, int liqflgsw, const double *cldfmcl,
                   const double *taucmcl, const double *ssacmcl, const double *asmcmcl,
                   const double *fsfcmcl, const double *ciwpmcl, const double *clwpmcl,
                   const double *reicmcl, const double *relqmcl, const double *tauaer,
                   const double *ssaaer, const double *asmaer, const double *ecaer, double *swuflx,
                   double *swdflx, double *swhr, double *swuflxc, double *swdflxc, double *swhrc,
                   const double *bndsolvar, const double *indsolvar, const double *solcycfrac,
                   void *stream) {
  (void)tlev, (void)tsfc, (void)n2ovmr;  // reach only unused setcoef locals / no taumol term
  if (int rc = check_init()) return rc;
  if (ncol < 0 || nlay < 4 || nlay > MAX_LAY)
    return fail(RRTMGSW_ERR_ARGUMENT, "nlay must lie in 4..256 and ncol >= 0");
  if (!icld || !iaer) return fail(RRTMGSW_ERR_ARGUMENT, "icld and iaer are required");
  if (*icld < 0 || *icld > 3) *icld = 2;                    // src/rrtmg_sw_rad.f90:589
  if (*iaer != 0 && *iaer != 6 && *iaer != 10) *iaer = 0;  // :597
  if (ncol == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  auto call = stream_call(s);
  RtArgs a{};
  a.ncol = ncol, a.nlay = nlay;
  a.play = play, a.plev = plev, a.tlay = tlay, a.h2ovmr = h2ovmr, a.o3vmr = o3vmr;
  a.co2vmr = co2vmr, a.ch4vmr = ch4vmr, a.o2vmr = o2vmr;
  a.asdir = asdir, a.asdif = asdif, a.aldir = aldir, a.aldif = aldif, a.coszen = coszen;
  a.cldfmcl = cldfmcl, a.taucmcl = taucmcl, a.ssacmcl = ssacmcl, a.asmcmcl = asmcmcl;
  a.fsfcmcl = fsfcmcl, a.ciwpmcl = ciwpmcl, a.clwpmcl = clwpmcl, a.reicmcl = reicmcl, a.relqmcl = relqmcl;
  a.tauaer = tauaer, a.ssaaer = ssaaer, a.asmaer = asmaer, a.ecaer = ecaer;
  a.swuflx = swuflx, a.swdflx = swdflx, a.swhr = swhr, a.swuflxc = swuflxc, a.swdflxc = swdflxc;
  a.swhrc = swhrc;
  a.fused = 0;
  solar_scalars(adjes, dyofyr, scon, isolvar, bndsolvar, indsolvar, solcycfrac, a.sc);
  a.sc.heatfac = ctx.heatfac;
  a.sc.icld = *icld, a.sc.iaer = *iaer;
  a.sc.inflag = inflgsw, a.sc.iceflag = iceflgsw, a.sc.liqflag = liqflgsw;
  const void *need[] = {play, plev, tlay, h2ovmr, o3vmr, co2vmr, ch4vmr, o2vmr, asdir, asdif,
                        aldir, aldif, coszen, swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc};
  for (const void *p : need)
    if (!p) return fail(RRTMGSW_ERR_ARGUMENT, "a required array argument is NULL");
  if (*icld >= 1 && (!cldfmcl || !taucmcl || !ssacmcl || !asmcmcl || !fsfcmcl || !ciwpmcl ||
                     !clwpmcl || !reicmcl || !relqmcl))
    return fail(RRTMGSW_ERR_ARGUMENT, "McICA cloud arrays are required when icld >= 1");
  if (*iaer == 10 && (!tauaer || !ssaaer || !asmaer))
    return fail(RRTMGSW_ERR_ARGUMENT, "tauaer, ssaaer and asmaer are required when iaer = 10");
  if (*iaer == 6 && !ecaer) return fail(RRTMGSW_ERR_ARGUMENT, "ecaer is required when iaer = 6");
  return run_rt(a, ctx.dev, s);
}

int rrtmgsw_sw(int ncol, int nlay, int *icld, int *iaer, const double *play, const double *plev,
               const double *tlay, const double *tlev, const double *tsfc, const double *h2ovmr,
               const double *o3vmr, const double *co2vmr, const double *ch4vmr,
               const double *n2ovmr, const double *o2vmr, const double *asdir, const double *asdif,
               const double *aldir, const double *aldif, const double *coszen, double adjes,
               int dyofyr, double scon, int isolvar, int inflgsw, int iceflgsw, int liqflgsw,
               const double *cldfmcl, const double *taucmcl, const double *ssacmcl,
               const double *asmcmcl, const double *fsfcmcl, const double *ciwpmcl,
               const double *clwpmcl, const double *reicmcl, const double *relqmcl,
               const double *tauaer, const double *ssaaer, const double *asmaer,
               const double *ecaer, double *swuflx, double *swdflx, double *swhr,
               double *swuflxc, double *swdflxc, double *swhrc, const double *bndsolvar,
               const double *indsolvar, const double *solcycfrac) {
  if (int rc = check_init()) return rc;
  if (ncol < 0 || nlay < 4 || nlay > MAX_LAY)
    return fail(RRTMGSW_ERR_ARGUMENT, "nlay must lie in 4..256 and ncol >= 0");
  if (!icld || !iaer) return fail(RRTMGSW_ERR_ARGUMENT, "icld and iaer are required");
  if (ncol == 0) {
    if (*icld < 0 || *icld > 3) *icld = 2;
    if (*iaer != 0 && *iaer != 6 && *iaer != 10) *iaer = 0;
    return 0;
  }
  const size_t C = ncol, L = nlay, G = NGPT;
  using In = HostIn;
  // the driver's clamps (src/rrtmg_sw_rad.f90:589, :597) decide which arrays are read
  int ic = (*icld < 0 || *icld > 3) ? 2 : *icld;
  int ia = (*iaer != 0 && *iaer != 6 && *iaer != 10) ? 0 : *iaer;
  bool clouds = ic >= 1;
  In in[] = {{play, C * L}, {plev, C * (L + 1)}, {tlay, C * L}, {h2ovmr, C * L}, {o3vmr, C * L},
             {co2vmr, C * L}, {ch4vmr, C * L}, {o2vmr, C * L}, {asdir, C}, {asdif, C}, {aldir, C},
             {aldif, C}, {coszen, C},
             {cldfmcl, clouds ? G * C * L : 0}, {taucmcl, clouds ? G * C * L : 0},
             {ssacmcl, clouds ? G * C * L : 0}, {asmcmcl, clouds ? G * C * L : 0},
             {fsfcmcl, clouds ? G * C * L : 0}, {ciwpmcl, clouds ? G * C * L : 0},
             {clwpmcl, clouds ? G * C * L : 0}, {reicmcl, clouds ? C * L : 0},
             {relqmcl, clouds ? C * L : 0},
             {tauaer, ia == 10 ? C * L * NBND : 0}, {ssaaer, ia == 10 ? C * L * NBND : 0},
             {asmaer, ia == 10 ? C * L * NBND : 0}, {ecaer, ia == 6 ? C * L * NAER : 0}};
  const int NIN = sizeof in / sizeof *in;
  const HostOut out[] = {{swuflx, C * (L + 1)}, {swdflx, C * (L + 1)}, {swuflxc, C * (L + 1)},
                         {swdflxc, C * (L + 1)}, {swhr, C * L}, {swhrc, C * L}};
  for (const HostOut &o : out)
    if (!o.h) return fail(RRTMGSW_ERR_ARGUMENT, "a required array argument is NULL");
  return host_call(in, NIN, out, 6, [&](const double *const *dp, double *const *o) {
    return rrtmgsw_sw_dev(ncol, nlay, icld, iaer, dp[0], dp[1], dp[2], nullptr, nullptr, dp[3], dp[4],
                          dp[5], dp[6], nullptr, dp[7], dp[8], dp[9], dp[10], dp[11], dp[12], adjes,
                          dyofyr, scon, isolvar, inflgsw, iceflgsw, liqflgsw, dp[13], dp[14], dp[15],
                          dp[16], dp[17], dp[18], dp[19], dp[20], dp[21], dp[22], dp[23], dp[24], dp[25],
                          o[0], o[1], o[4], o[2], o[3], o[5], bndsolvar, indsolvar, solcycfrac, nullptr);
  });
}

// Non-McICA driver (src/rrtmg_sw_rad.nomcica.f90:97-816): layer cloud fractions and cldprop_sw
// optics; inputs already on the device.
int rrtmgsw_sw_nomcica_dev(int ncol, int nlay, int *icld, int *iaer, const double *play,
                           const double *plev, const double *tlay, const double *tlev,
                           const double *tsfc, const double *h2ovmr, const double *o3vmr,
                           const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                           const double *o2vmr, const double *asdir, const double *asdif,
                           const double *aldir, const double *aldif, const double *coszen,
                           double adjes, int dyofyr, double scon, int isolvar, int inflgsw,
                           int iceflgsw, int liqflgsw, const double *cldfr, const double *taucld,
                           const double *ssacld, const double *asmcld, const double *fsfcld,
                           const double *cicewp, const double *cliqwp, const double *reice,
                           const double *reliq, const double *tauaer, const double *ssaaer,
                           const double *asmaer, const double *ecaer, double *swuflx, double *swdflx,
                           double *swhr, double *swuflxc, double *swdflxc, double *swhrc,
                           const double *bndsolvar, const double *indsolvar,
                           const double *solcycfrac, void *stream) {
  (void)tlev, (void)tsfc, (void)n2ovmr;  // reach only unused setcoef locals / no taumol term
  if (int rc = check_init()) return rc;
  if (ncol < 0 || nlay < 4 || nlay > MAX_LAY)
    return fail(RRTMGSW_ERR_ARGUMENT, "nlay must lie in 4..256 and ncol >= 0");
  if (!icld || !iaer) return fail(RRTMGSW_ERR_ARGUMENT, "icld and iaer are required");
  if (*icld < 0 || *icld > 3) *icld = 2;                    // src/rrtmg_sw_rad.nomcica.f90:560
  if (*iaer != 0 && *iaer != 6 && *iaer != 10) *iaer = 0;  // :568
  if (ncol == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  auto call = stream_call(s);
  RtArgs a{};
  a.ncol = ncol, a.nlay = nlay;
  a.play = play, a.plev = plev, a.tlay = tlay, a.h2ovmr = h2ovmr, a.o3vmr = o3vmr;
  a.co2vmr = co2vmr, a.ch4vmr = ch4vmr, a.o2vmr = o2vmr;
  a.asdir = asdir, a.asdif = asdif, a.aldir = aldir, a.aldif = aldif, a.coszen = coszen;
  a.tauaer = tauaer, a.ssaaer = ssaaer, a.asmaer = asmaer, a.ecaer = ecaer;
  a.swuflx = swuflx, a.swdflx = swdflx, a.swhr = swhr, a.swuflxc = swuflxc, a.swdflxc = swdflxc;
  a.swhrc = swhrc;
  a.fused = 0, a.layer_clouds = 1;
  a.cldfrac = cldfr, a.ciwp = cicewp, a.clwp = cliqwp, a.rei = reice, a.rel = reliq;
  a.tauc = taucld, a.ssac = ssacld, a.asmc = asmcld, a.fsfc = fsfcld;
  solar_scalars(adjes, dyofyr, scon, isolvar, bndsolvar, indsolvar, solcycfrac, a.sc);
  a.sc.heatfac = ctx.heatfac;
  a.sc.icld = *icld, a.sc.iaer = *iaer;
  a.sc.inflag = inflgsw, a.sc.iceflag = iceflgsw, a.sc.liqflag = liqflgsw;
  const void *need[] = {play, plev, tlay, h2ovmr, o3vmr, co2vmr, ch4vmr, o2vmr, asdir, asdif,
                        aldir, aldif, coszen, swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc};
  for (const void *p : need)
    if (!p) return fail(RRTMGSW_ERR_ARGUMENT, "a required array argument is NULL");
  if (*icld >= 1 && (!cldfr || !taucld || !ssacld || !asmcld || !fsfcld || !cicewp || !cliqwp ||
                     !reice || !reliq))
    return fail(RRTMGSW_ERR_ARGUMENT, "cloud arrays are required when icld >= 1");
  if (*iaer == 10 && (!tauaer || !ssaaer || !asmaer))
    return fail(RRTMGSW_ERR_ARGUMENT, "tauaer, ssaaer and asmaer are required when iaer = 10");
  if (*iaer == 6 && !ecaer) return fail(RRTMGSW_ERR_ARGUMENT, "ecaer is required when iaer = 6");
  return run_rt(a, ctx.dev, s);
}

// Host arrays in and out of the non-McICA driver: one device allocation for the inputs the
// flags select and the outputs, synchronous copies, the device entry on the null stream.
int rrtmgsw_sw_nomcica(int ncol, int nlay, int *icld, int *iaer, const double *play,
                       const double *plev, const double *tlay, const double *tlev,
                       const double *tsfc, const double *h2ovmr, const double *o3vmr,
                       const double *co2vmr, const double *ch4vmr, const double *n2ovmr,
                       const double *o2vmr, const double *asdir, const double *asdif,
                       const double *aldir, const double *aldif, const double *coszen,
                       double adjes, int dyofyr, double scon, int isolvar, int inflgsw,
                       int iceflgsw, int liqflgsw, const double *cldfr, const double *taucld,
                       const double *ssacld, const double *asmcld, const double *fsfcld,
                       const double *cicewp, const double *cliqwp, const double *reice,
                       const double *reliq, const double *tauaer, const double *ssaaer,
                       const double *asmaer, const double *ecaer, double *swuflx, double *swdflx,
                       double *swhr, double *swuflxc, double *swdflxc, double *swhrc,
                       const double *bndsolvar, const double *indsolvar, const double *solcycfrac) {
  if (int rc = check_init()) return rc;
  if (ncol < 0 || nlay < 4 || nlay > MAX_LAY)
    return fail(RRTMGSW_ERR_ARGUMENT, "nlay must lie in 4..256 and ncol >= 0");
  if (!icld || !iaer) return fail(RRTMGSW_ERR_ARGUMENT, "icld and iaer are required");
  if (ncol == 0) {
    if (*icld < 0 || *icld > 3) *icld = 2;
    if (*iaer != 0 && *iaer != 6 && *iaer != 10) *iaer = 0;
    return 0;
  }
  const size_t C = ncol, L = nlay;
  using In = HostIn;
  int ic = (*icld < 0 || *icld > 3) ? 2 : *icld;
  int ia = (*iaer != 0 && *iaer != 6 && *iaer != 10) ? 0 : *iaer;
  const size_t cl = ic >= 1 ? C * L : 0, cb = ic >= 1 ? NBND * C * L : 0;
  In in[] = {{play, C * L}, {plev, C * (L + 1)}, {tlay, C * L}, {h2ovmr, C * L}, {o3vmr, C * L},
             {co2vmr, C * L}, {ch4vmr, C * L}, {o2vmr, C * L}, {asdir, C}, {asdif, C}, {aldir, C},
             {aldif, C}, {coszen, C}, {cldfr, cl}, {taucld, cb}, {ssacld, cb}, {asmcld, cb},
             {fsfcld, cb}, {cicewp, cl}, {cliqwp, cl}, {reice, cl}, {reliq, cl},
             {tauaer, ia == 10 ? C * L * NBND : 0}, {ssaaer, ia == 10 ? C * L * NBND : 0},
             {asmaer, ia == 10 ? C * L * NBND : 0}, {ecaer, ia == 6 ? C * L * NAER : 0}};
  const int NIN = sizeof in / sizeof *in;
  const HostOut out[] = {{swuflx, C * (L + 1)}, {swdflx, C * (L + 1)}, {swuflxc, C * (L + 1)},
                         {swdflxc, C * (L + 1)}, {swhr, C * L}, {swhrc, C * L}};
  for (const HostOut &o : out)
    if (!o.h) return fail(RRTMGSW_ERR_ARGUMENT, "a required array argument is NULL");
  return host_call(in, NIN, out, 6, [&](const double *const *dp, double *const *o) {
    return rrtmgsw_sw_nomcica_dev(ncol, nlay, icld, iaer, dp[0], dp[1], dp[2], nullptr, nullptr, dp[3],
                                  dp[4], dp[5], dp[6], nullptr, dp[7], dp[8], dp[9], dp[10], dp[11],
                                  dp[12], adjes, dyofyr, scon, isolvar, inflgsw, iceflgsw, liqflgsw,
                                  dp[13], dp[14], dp[15], dp[16], dp[17], dp[18], dp[19], dp[20], dp[21],
                                  dp[22], dp[23], dp[24], dp[25], o[0], o[1], o[4], o[2], o[3], o[5],
                                  bndsolvar, indsolvar, solcycfrac, nullptr);
  });
}

// Column program without McICA (src/rrtmg_sw.1col.f90:540-1000 for imca = 0): layer amounts in,
// broadband and per-band up / down / direct-down fluxes out. Host arrays.
int rrtmgsw_column_sw(int ncol, int nlay, int icld, int idelm, const double *pavel,
                      const double *tavel, const double *pz, const double *coldry,
                      const double *wkl, const double *albedo, const double *coszen, double adjes,
                      int dyofyr, double scon, int isolvar, const double *bndsolvar,
                      const double *indsolvar, const double *solcycfrac, int inflag, int iceflag,
                      int liqflag, const double *cldfrac, const double *ciwp, const double *clwp,
                      const double *rei, const double *rel, const double *tauc, const double *ssac,
                      const double *asmc, const double *fsfc, const double *tauaer,
                      const double *ssaaer, const double *asmaer, double *fup, double *fdn,
                      double *fdir) {
  if (int rc = check_init()) return rc;
  auto call = stream_call(nullptr);
  if (ncol < 0 || nlay < 4 || nlay > MAX_LAY)
    return fail(RRTMGSW_ERR_ARGUMENT, "nlay must lie in 4..256 and ncol >= 0");
  if (idelm != 0 && idelm != 1) return fail(RRTMGSW_ERR_ARGUMENT, "idelm must be 0 or 1");
  if (ncol == 0) return 0;
  const void *need[] = {pavel, tavel, pz, coldry, wkl, albedo, coszen, fup, fdn, fdir};
  for (const void *p : need)
    if (!p) return fail(RRTMGSW_ERR_ARGUMENT, "a required array argument is NULL");
  const bool clouds = icld >= 1, aer = tauaer != nullptr;
  if (clouds && (!cldfrac || !ciwp || !clwp || !rei || !rel || !tauc || !ssac || !asmc || !fsfc))
    return fail(RRTMGSW_ERR_ARGUMENT, "cloud arrays are required when icld >= 1");
  if (aer && (!ssaaer || !asmaer))
    return fail(RRTMGSW_ERR_ARGUMENT, "tauaer, ssaaer and asmaer go together");
  const size_t C = ncol, L = nlay;
  // the column program stops on any fraction other than 0 and 1 (:612-616)
  if (clouds)
    for (size_t q = 0; q < C * L; q++)
      if (cldfrac[q] != 0.0 && cldfrac[q] != 1.0)
        return fail(RRTMGSW_ERR_PARTIAL_CLOUD, "PARTIAL CLOUD NOT ALLOWED FOR IMCA=0");
  struct In {
    const double *h;
    size_t n;
  };
  In in[] = {{pavel, C * L}, {tavel, C * L}, {pz, C * (L + 1)}, {coldry, C * L}, {wkl, 7 * C * L},
             {albedo, NBND * C}, {coszen, C}, {cldfrac, clouds ? C * L : 0},
             {ciwp, clouds ? C * L : 0}, {clwp, clouds ? C * L : 0}, {rei, clouds ? C * L : 0},
             {rel, clouds ? C * L : 0}, {tauc, clouds ? NBND * C * L : 0},
             {ssac, clouds ? NBND * C * L : 0}, {asmc, clouds ? NBND * C * L : 0},
             {fsfc, clouds ? NBND * C * L : 0}, {tauaer, aer ? C * L * NBND : 0},
             {ssaaer, aer ? C * L * NBND : 0}, {asmaer, aer ? C * L * NBND : 0}};
  const int NIN = sizeof in / sizeof *in;
  size_t total = 0;
  for (auto &x : in) total += x.n;
  const size_t nlev = C * (L + 1), nband = nlev * NBND;
  const size_t nout = 4 * nlev + 2 * C * L + 3 * nband;
  double *d = nullptr;
  HIPCHK(hipMalloc(&d, (total + nout) * sizeof(double)));
  const double *dp[NIN];
  size_t off = 0;
  for (int i = 0; i < NIN; i++) {
    dp[i] = in[i].n ? d + off : nullptr;
    if (in[i].n) {
      hipError_t e = hipMemcpy(d + off, in[i].h, in[i].n * sizeof(double), hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        hipFree(d);
        return fail(RRTMGSW_ERR_HIP, hipGetErrorString(e));
      }
    }
    off += in[i].n;
  }
  double *o = d + total;
  RtArgs a{};
  a.ncol = ncol, a.nlay = nlay;
  a.play = dp[0], a.tlay = dp[1], a.plev = dp[2], a.coldry_in = dp[3], a.wkl_in = dp[4];
  a.albbnd = dp[5], a.coszen = dp[6];
  a.cldfrac = dp[7], a.ciwp = dp[8], a.clwp = dp[9], a.rei = dp[10], a.rel = dp[11];
  a.tauc = dp[12], a.ssac = dp[13], a.asmc = dp[14], a.fsfc = dp[15];
  a.tauaer = dp[16], a.ssaaer = dp[17], a.asmaer = dp[18];
  a.swuflx = o, a.swdflx = o + nlev, a.swuflxc = o + 2 * nlev, a.swdflxc = o + 3 * nlev;
  a.swhr = o + 4 * nlev, a.swhrc = a.swhr + C * L;
  a.bandu = a.swhrc + C * L, a.bandd = a.bandu + nband, a.banddir = a.bandd + nband;
  a.fused = 0, a.layer_clouds = 1, a.idelm = idelm;
  solar_scalars(adjes, dyofyr, scon, isolvar, bndsolvar, indsolvar, solcycfrac, a.sc);
  a.sc.heatfac = ctx.heatfac;
  a.sc.icld = clouds ? 1 : 0, a.sc.iaer = aer ? 10 : 0;
  a.sc.inflag = inflag, a.sc.iceflag = iceflag, a.sc.liqflag = liqflag;
  int rc = run_rt(a, ctx.dev, nullptr, RRTMGSW_FP64, true);
  if (rc == 0) {
    std::vector<double> h(nout);
    hipError_t e = hipMemcpy(h.data(), o, nout * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      rc = fail(RRTMGSW_ERR_HIP, hipGetErrorString(e));
    } else {
      const double *bu = h.data() + 4 * nlev + 2 * C * L, *bd = bu + nband, *bdir = bd + nband;
      for (size_t q = 0; q < nlev; q++) {
        double dsum = 0.0;
        for (int b = 0; b < NBND; b++) {
          fup[q + nlev * (b + 1)] = bu[q + nlev * b];
          fdn[q + nlev * (b + 1)] = bd[q + nlev * b];
          fdir[q + nlev * (b + 1)] = bdir[q + nlev * b];
          dsum += bdir[q + nlev * b];
        }
        fup[q] = h[q], fdn[q] = h[nlev + q], fdir[q] = dsum;
      }
    }
  }
  hipFree(d);
  return rc;
}

// The column program with McICA (src/rrtmg_sw.1col.f90:577-600 with imca = 1): nsamples draws
// of one column, sample s (ims = s + 1) with permuteseed (ims + 1) * 112 as
// src/mcica_subcol_gen_sw.1col.f90 sets it, all samples in one launch (a workgroup each).
// Outputs per sample, (nsamples, nlay+1, 15) as rrtmgsw_column_sw's.
int rrtmgsw_column_mcica_sw(int nlay, int nsamples, int icld, int irng, int idelm,
                            const double *pavel, const double *tavel, const double *pz,
                            const double *coldry, const double *wkl, const double *albedo,
                            double coszen, double adjes, int dyofyr, double scon, int isolvar,
                            const double *bndsolvar, const double *indsolvar,
                            const double *solcycfrac, int inflag, int iceflag, int liqflag,
                            const double *cldfrac, const double *ciwp, const double *clwp,
                            const double *rei, const double *rel, const double *tauc,
                            const double *ssac, const double *asmc, const double *fsfc,
                            const double *alpha, const double *tauaer, const double *ssaaer,
                            const double *asmaer, double *fup, double *fdn, double *fdir) {
  if (int rc = check_init()) return rc;
  auto call = stream_call(nullptr);
  if (nsamples < 0 || nlay < 4 || nlay > MAX_LAY)
    return fail(RRTMGSW_ERR_ARGUMENT, "nlay must lie in 4..256 and nsamples >= 0");
  if (icld < 1 || icld > 5) return fail(RRTMGSW_ERR_ICLD, "RRTMG_SW: INVALID ICLD");
  if (idelm != 0 && idelm != 1) return fail(RRTMGSW_ERR_ARGUMENT, "idelm must be 0 or 1");
  if (nsamples == 0) return 0;
  const void *need[] = {pavel, tavel, pz, coldry, wkl, albedo, cldfrac, ciwp, clwp, rei, rel,
                        tauc, ssac, asmc, fsfc, fup, fdn, fdir};
  for (const void *p : need)
    if (!p) return fail(RRTMGSW_ERR_ARGUMENT, "a required array argument is NULL");
  if (icld >= 4 && !alpha) return fail(RRTMGSW_ERR_ARGUMENT, "alpha is required when icld = 4 or 5");
  const bool aer = tauaer != nullptr;
  if (aer && (!ssaaer || !asmaer)) return fail(RRTMGSW_ERR_ARGUMENT, "tauaer, ssaaer and asmaer go together");
  const int ir = irng != 0 ? 1 : 0;
  const size_t C = nsamples, L = nlay;
  // every sample is the same column: (ncol, ...) layouts with the column repeated
  std::vector<double> h;
  auto rep2 = [&](const double *x, size_t inner, size_t outer) {   // x (inner, outer) -> (inner, C, outer)
    size_t off = h.size();
    h.resize(off + inner * C * outer);
    for (size_t o = 0; o < outer; o++)
      for (size_t c = 0; c < C; c++)
        for (size_t i = 0; i < inner; i++) h[off + i + inner * (c + C * o)] = x[i + inner * o];
    return off;
  };
  size_t o_pavel = rep2(pavel, 1, L), o_tavel = rep2(tavel, 1, L), o_pz = rep2(pz, 1, L + 1);
  size_t o_coldry = rep2(coldry, 1, L), o_wkl = rep2(wkl, 7, L), o_alb = rep2(albedo, NBND, 1);
  std::vector<double> mu(C, coszen);
  size_t o_mu = h.size();
  h.insert(h.end(), mu.begin(), mu.end());
  size_t o_cf = rep2(cldfrac, 1, L), o_ci = rep2(ciwp, 1, L), o_cl = rep2(clwp, 1, L);
  size_t o_rei = rep2(rei, 1, L), o_rel = rep2(rel, 1, L), o_tc = rep2(tauc, NBND, L);
  size_t o_sc = rep2(ssac, NBND, L), o_ac = rep2(asmc, NBND, L), o_fc = rep2(fsfc, NBND, L);
  size_t o_al = icld >= 4 ? rep2(alpha, 1, L) : 0;
  size_t o_ta = 0, o_sa = 0, o_ga = 0;
  if (aer) {   // (nlay, 14) -> (C, nlay, 14)
    o_ta = rep2(tauaer, 1, L * NBND), o_sa = rep2(ssaaer, 1, L * NBND), o_ga = rep2(asmaer, 1, L * NBND);
  }
  const size_t nlev = C * (L + 1), nband = nlev * NBND;
  const size_t nout = 4 * nlev + 2 * C * L + 3 * nband;
  // KISS: one jump per sample (its permuteseed) ahead of the per-sub-column jumps;
  // MT19937: one stream per sample, seeded with its permuteseed
  const int per_sub = icld == 3 ? 1 : (icld >= 4 ? 2 * nlay : nlay);
  const size_t ndraw = (size_t)NGPT * (icld == 3 ? 1 : (size_t)per_sub);
  std::vector<KissJump> cj;
  if (ir == 0)
    for (size_t c = 0; c < C; c++) cj.push_back(kiss_jump_of((uint64_t)(c + 2) * NGPT));
  double *d = nullptr;
  const size_t bytes = (h.size() + nout) * sizeof(double) + cj.size() * sizeof(KissJump) +
                       (ir ? C * ndraw * 4 : 0) + 256;
  HIPCHK(hipMalloc(&d, bytes));
  hipError_t e = hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
  double *o = d + h.size();
  char *tail = reinterpret_cast<char *>(o + nout);
  uint32_t *d_cj = reinterpret_cast<uint32_t *>(tail);
  uint32_t *d_mt = reinterpret_cast<uint32_t *>(tail + cj.size() * sizeof(KissJump));
  if (e == hipSuccess && !cj.empty())
    e = hipMemcpy(d_cj, cj.data(), cj.size() * sizeof(KissJump), hipMemcpyHostToDevice);
  if (e == hipSuccess && ir)
    for (size_t c = 0; c < C && e == hipSuccess; c++)
      e = launch_mt_stream((uint32_t)((c + 2) * NGPT), (long long)ndraw, d_mt + c * ndraw, nullptr, nullptr);
  if (e != hipSuccess) {
    hipFree(d);
    return fail(RRTMGSW_ERR_HIP, hipGetErrorString(e));
  }
  const uint32_t *jumps = nullptr;
  if (ir == 0)
    if (int rc = kiss_jumps(0, per_sub, nullptr, &jumps)) {
      hipFree(d);
      return rc;
    }
  RtArgs a{};
  a.ncol = (int)C, a.nlay = nlay;
  a.play = d + o_pavel, a.tlay = d + o_tavel, a.plev = d + o_pz, a.coldry_in = d + o_coldry;
  a.wkl_in = d + o_wkl, a.albbnd = d + o_alb, a.coszen = d + o_mu;
  a.cldfrac = d + o_cf, a.ciwp = d + o_ci, a.clwp = d + o_cl, a.rei = d + o_rei, a.rel = d + o_rel;
  a.tauc = d + o_tc, a.ssac = d + o_sc, a.asmc = d + o_ac, a.fsfc = d + o_fc;
  a.alpha = icld >= 4 ? d + o_al : nullptr;
  if (aer) a.tauaer = d + o_ta, a.ssaaer = d + o_sa, a.asmaer = d + o_ga;
  a.swuflx = o, a.swdflx = o + nlev, a.swuflxc = o + 2 * nlev, a.swdflxc = o + 3 * nlev;
  a.swhr = o + 4 * nlev, a.swhrc = a.swhr + C * L;
  a.bandu = a.swhrc + C * L, a.bandd = a.bandu + nband, a.banddir = a.bandd + nband;
  a.fused = 1, a.mc_icld = icld, a.mc_irng = ir, a.permuteseed = 0, a.idelm = idelm;
  a.col_jump = ir == 0 ? d_cj : nullptr;
  a.mt_stream = ir ? d_mt : nullptr;
  a.mt_stride = ir ? ndraw : 0;
  DeviceTables t = ctx.dev;
  t.kiss_jump = jumps;
  solar_scalars(adjes, dyofyr, scon, isolvar, bndsolvar, indsolvar, solcycfrac, a.sc);
  a.sc.heatfac = ctx.heatfac;
  a.sc.icld = 1, a.sc.iaer = aer ? 10 : 0;
  a.sc.inflag = inflag, a.sc.iceflag = iceflag, a.sc.liqflag = liqflag;
  int rc = run_rt(a, t, nullptr, RRTMGSW_FP64, true);
  if (rc == 0) {
    std::vector<double> r(nout);
    e = hipMemcpy(r.data(), o, nout * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      rc = fail(RRTMGSW_ERR_HIP, hipGetErrorString(e));
    } else {
      const double *bu = r.data() + 4 * nlev + 2 * C * L, *bd = bu + nband, *bdir = bd + nband;
      for (size_t q = 0; q < nlev; q++) {
        double dsum = 0.0;
        for (int b = 0; b < NBND; b++) {
          fup[q + nlev * (b + 1)] = bu[q + nlev * b];
          fdn[q + nlev * (b + 1)] = bd[q + nlev * b];
          fdir[q + nlev * (b + 1)] = bdir[q + nlev * b];
          dsum += bdir[q + nlev * b];
        }
        fup[q] = r[q], fdn[q] = r[nlev + q], fdir[q] = dsum;
      }
    }
  }
  hipFree(d);
  return rc;
}

static int mcica_prepare(int ncol, int nlay, int icld, int permuteseed, int irng, hipStream_t s,
                         const uint32_t **mt, const uint32_t **jumps) {
  *mt = nullptr;
  *jumps = nullptr;
  int per_sub = icld == 3 ? 1 : (icld >= 4 ? 2 * nlay : nlay);
  if (irng == 0) return kiss_jumps(permuteseed, per_sub, s, jumps);
  long long ndraw = (long long)NGPT * ncol * (icld == 3 ? 1 : (icld >= 4 ? 2LL * nlay : nlay));
  void *buf = nullptr;
  MtJumpTable jt{};
  const char *seq = std::getenv("RRTMGSW_MT_SEQUENTIAL");   // 1: one workgroup, no jump-ahead
  const bool jump = mt_extra_words(ndraw) > 0 && !(seq && std::atoi(seq) != 0);
  {
    std::lock_guard<std::mutex> lk(ctx.mu);
    if (jump) {
      const int lj = mt_log2_seg(ndraw);
      const long long J = 1LL << lj;
      const int need = (int)((ndraw + J - 1) / J) - 1;
      Context::MtJump &t = ctx.mtjump[lj];
      if (t.count < need) {
        // x^(k J) mod P for every segment start k, as lists of set bits with offsets
        const int count = std::max(need, 2 * t.count);
        std::vector<uint32_t> c = mt_jump_multiples(lj, count);
        if (c.size() != (size_t)count * 624)
          return fail(RRTMGSW_ERR_ARGUMENT, "MT19937 jump-ahead polynomial not found");
        std::vector<uint32_t> bits, offs(1, 0);
        for (int k = 0; k < count; k++) {
          const uint32_t *w = c.data() + (size_t)k * 624;
          for (uint32_t i = 0; i < 19937; i++)
            if ((w[i >> 5] >> (i & 31)) & 1u) bits.push_back(i);
          offs.push_back((uint32_t)bits.size());
        }
        if (t.bits) ctx.retired.push_back(t.bits), ctx.retired.push_back(t.offs);
        t.bits = t.offs = nullptr, t.count = 0;
        HIPCHK(hipMalloc(&t.bits, bits.size() * sizeof(uint32_t)));
        HIPCHK(hipMalloc(&t.offs, offs.size() * sizeof(uint32_t)));
        HIPCHK(hipMemcpy(t.bits, bits.data(), bits.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(t.offs, offs.data(), offs.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        t.count = count;
      }
      jt.log2_seg = lj, jt.count = t.count, jt.bits = t.bits, jt.offs = t.offs;
    }
    const size_t words = (size_t)ndraw + (jump ? mt_extra_words(ndraw) : 0);
    if (int rc = stream_buffer(ctx.mtbuf, s, words * 4, &buf)) return rc;
  }
  HIPCHK(launch_mt_stream((uint32_t)permuteseed, ndraw, (uint32_t *)buf, jump ? &jt : nullptr, s));
  *mt = (const uint32_t *)buf;
  return 0;
}

int rrtmgsw_mcica_subcol_dev(int iplon, int ncol, int nlay, int icld, int permuteseed, int *irng,
                             const double *play, const double *cldfrac, const double *ciwp,
                             const double *clwp, const double *rei, const double *rel,
                             const double *tauc, const double *ssac, const double *asmc,
                             const double *fsfc, const double *alpha, double *cldfmcl,
                             double *ciwpmcl, double *clwpmcl, double *reicmcl, double *relqmcl,
                             double *taucmcl, double *ssacmcl, double *asmcmcl, double *fsfcmcl,
                             void *stream) {
  (void)iplon;
  if (int rc = check_init()) return rc;
  if (icld == 0) return 0;  // src/mcica_subcol_gen_sw.f90:261
  if (icld < 0 || icld > 5) return fail(RRTMGSW_ERR_ICLD, "MCICA_SUBCOL: INVALID ICLD");
  if (!irng) return fail(RRTMGSW_ERR_ARGUMENT, "irng is required");
  if (*irng != 0) *irng = 1;  // :450
  if (ncol <= 0) return 0;
  if (*irng == 0 && nlay < 4) return fail(RRTMGSW_ERR_ARGUMENT, "KISS seeding needs at least four layers");
  if (nlay > MAX_LAY) return fail(RRTMGSW_ERR_ARGUMENT, "nlay must not exceed 256");
  const void *need[] = {play, cldfrac, ciwp, clwp, rei, rel, tauc, ssac, asmc, fsfc, cldfmcl, ciwpmcl,
                        clwpmcl, reicmcl, relqmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl};
  for (const void *p : need)
    if (!p) return fail(RRTMGSW_ERR_ARGUMENT, "a required array argument is NULL");
  if (icld >= 4 && !alpha) return fail(RRTMGSW_ERR_ARGUMENT, "alpha is required when icld = 4 or 5");
  hipStream_t s = (hipStream_t)stream;
  auto call = stream_call(s);
  const uint32_t *mt, *jumps;
  if (int rc = mcica_prepare(ncol, nlay, icld, permuteseed, *irng, s, &mt, &jumps)) return rc;
  void *st = nullptr;
  {
    std::lock_guard<std::mutex> lk(ctx.mu);
    if (int rc = stream_buffer(ctx.status, s, 256, &st)) return rc;
  }
  hipLaunchKernelGGL(zero_int, dim3(1), dim3(1), 0, s, (int *)st);
  McicaArgs a{};
  a.ncol = ncol, a.nlay = nlay, a.icld = icld, a.irng = *irng, a.permuteseed = permuteseed;
  a.play = play, a.cldfrac = cldfrac, a.ciwp = ciwp, a.clwp = clwp, a.rei = rei, a.rel = rel;
  a.tauc = tauc, a.ssac = ssac, a.asmc = asmc, a.fsfc = fsfc, a.alpha = alpha, a.mt_stream = mt;
  a.cldfmcl = cldfmcl, a.ciwpmcl = ciwpmcl, a.clwpmcl = clwpmcl, a.reicmcl = reicmcl;
  a.relqmcl = relqmcl, a.taucmcl = taucmcl, a.ssacmcl = ssacmcl, a.asmcmcl = asmcmcl, a.fsfcmcl = fsfcmcl;
  a.status = (int *)st;
  DeviceTables t = ctx.dev;
  t.kiss_jump = jumps;
  HIPCHK(launch_mcica(a, t, s));
  return finish(s, (int *)st);
}

int rrtmgsw_mcica_subcol(int iplon, int ncol, int nlay, int icld, int permuteseed, int *irng,
                         const double *play, const double *cldfrac, const double *ciwp,
                         const double *clwp, const double *rei, const double *rel,
                         const double *tauc, const double *ssac, const double *asmc,
                         const double *fsfc, const double *alpha, double *cldfmcl,
                         double *ciwpmcl, double *clwpmcl, double *reicmcl, double *relqmcl,
                         double *taucmcl, double *ssacmcl, double *asmcmcl, double *fsfcmcl) {
  if (int rc = check_init()) return rc;
  if (icld == 0) return 0;
  if (icld < 0 || icld > 5) return fail(RRTMGSW_ERR_ICLD, "MCICA_SUBCOL: INVALID ICLD");
  if (ncol <= 0) {
    if (irng && *irng != 0) *irng = 1;
    return 0;
  }
  const size_t C = ncol, L = nlay, G = NGPT;
  const HostIn in[] = {{play, C * L}, {cldfrac, C * L}, {ciwp, C * L}, {clwp, C * L}, {rei, C * L},
                       {rel, C * L}, {tauc, NBND * C * L}, {ssac, NBND * C * L}, {asmc, NBND * C * L},
                       {fsfc, NBND * C * L}, {alpha, icld >= 4 ? C * L : 0}};
  const HostOut out[] = {{cldfmcl, G * C * L}, {ciwpmcl, G * C * L}, {clwpmcl, G * C * L},
                         {reicmcl, C * L}, {relqmcl, C * L}, {taucmcl, G * C * L},
                         {ssacmcl, G * C * L}, {asmcmcl, G * C * L}, {fsfcmcl, G * C * L}};
  for (const HostOut &o : out)
    if (!o.h) return fail(RRTMGSW_ERR_ARGUMENT, "a required array argument is NULL");
  return host_call(in, 11, out, 9, [&](const double *const *dp, double *const *o) {
    return rrtmgsw_mcica_subcol_dev(iplon, ncol, nlay, icld, permuteseed, irng, dp[0], dp[1], dp[2],
                                    dp[3], dp[4], dp[5], dp[6], dp[7], dp[8], dp[9], dp[10], o[0], o[1],
                                    o[2], o[3], o[4], o[5], o[6], o[7], o[8], nullptr);
  });
}

int rrtmgsw_get_alpha(int iplon, int ncol, int nlayers, int icld, int idcor, double decorr_con,
                      const double *dz, const double *lat, int juldat, const double *cldfrac,
                      double *alpha) {
  (void)iplon;
  // src/mcica_subcol_gen_sw.f90:110-176 (default-real constants as double, the -r8 build)
  const double am1 = 1.4315, am2 = 2.1219, am4 = -25.584, amr = 7.0;
  if (!(icld == 4 || icld == 5)) return 0;
  for (int i = 0; i < ncol; i++) {
    double len = decorr_con;
    if (idcor == 1) {
      double am3 = juldat > 181 ? -4. * amr / 365. * (juldat - 272) : 4. * amr / 365. * (juldat - 91);
      double d = lat[i] - am3;
      len = (am1 + am2 * exp(-(d * d) / (am4 * am4))) * 1.e3;
    }
    double inv = len >= 0.0 ? 1.0 / len : 1.0;
    alpha[i] = 0.0;
    for (int k = 1; k < nlayers; k++) {
      size_t q = i + (size_t)ncol * k, qm = i + (size_t)ncol * (k - 1);
      alpha[q] = exp(-(0.5 * (dz[q] + dz[qm])) * inv);
      if (icld == 5 && cldfrac[q] == 0.0 && cldfrac[qm] > 0.0) alpha[q] = 0.0;
    }
  }
  return 0;
}

// Inputs, flags and McICA generator state of the fused entry points (mcica_subcol_sw followed by
// rrtmg_sw): validation as the two reference routines do it, clamps, solar scalars.
static int fused_setup(int ncol, int nlay, int icld, int iaer, int permuteseed, int *irng,
                       const double *play, const double *plev, const double *tlay,
                       const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                       const double *ch4vmr, const double *o2vmr, const double *asdir,
                       const double *asdif, const double *aldir, const double *aldif,
                       const double *coszen, double adjes, int dyofyr, double scon, int isolvar,
                       int inflgsw, int iceflgsw, int liqflgsw, const double *cldfrac,
                       const double *ciwp, const double *clwp, const double *rei,
                       const double *rel, const double *tauc, const double *ssac,
                       const double *asmc, const double *fsfc, const double *alpha,
                       const double *tauaer, const double *ssaaer, const double *asmaer,
                       const double *ecaer, const double *bndsolvar, const double *indsolvar,
                       const double *solcycfrac, hipStream_t s, RtArgs &a, DeviceTables &t,
                       bool prepare = true) {
  if (icld < 0 || icld > 5) return fail(RRTMGSW_ERR_ICLD, "MCICA_SUBCOL: INVALID ICLD");
  if (ncol < 0 || nlay < 4 || nlay > MAX_LAY)
    return fail(RRTMGSW_ERR_ARGUMENT, "nlay must lie in 4..256 and ncol >= 0");
  int ir = irng ? *irng : 0;
  if (ir != 0) ir = 1;
  if (irng) *irng = ir;
  if (iaer != 0 && iaer != 6 && iaer != 10) iaer = 0;
  a.ncol = ncol, a.nlay = nlay;
  a.play = play, a.plev = plev, a.tlay = tlay, a.h2ovmr = h2ovmr, a.o3vmr = o3vmr;
  a.co2vmr = co2vmr, a.ch4vmr = ch4vmr, a.o2vmr = o2vmr;
  a.asdir = asdir, a.asdif = asdif, a.aldir = aldir, a.aldif = aldif, a.coszen = coszen;
  a.tauaer = tauaer, a.ssaaer = ssaaer, a.asmaer = asmaer, a.ecaer = ecaer;
  a.fused = 1, a.mc_icld = icld, a.mc_irng = ir, a.permuteseed = permuteseed;
  a.cldfrac = cldfrac, a.ciwp = ciwp, a.clwp = clwp, a.rei = rei, a.rel = rel;
  a.tauc = tauc, a.ssac = ssac, a.asmc = asmc, a.fsfc = fsfc, a.alpha = alpha;
  solar_scalars(adjes, dyofyr, scon, isolvar, bndsolvar, indsolvar, solcycfrac, a.sc);
  a.sc.heatfac = ctx.heatfac;
  a.sc.icld = icld >= 1 ? 1 : 0, a.sc.iaer = iaer;
  a.sc.inflag = inflgsw, a.sc.iceflag = iceflgsw, a.sc.liqflag = liqflgsw;
  if (ncol == 0) return 0;
  const void *need[] = {play, plev, tlay, h2ovmr, o3vmr, co2vmr, ch4vmr, o2vmr, asdir, asdif,
                        aldir, aldif, coszen};
  for (const void *p : need)
    if (!p) return fail(RRTMGSW_ERR_ARGUMENT, "a required array argument is NULL");
  if (icld >= 1 && (!cldfrac || !ciwp || !clwp || !rei || !rel || !tauc || !ssac || !asmc || !fsfc))
    return fail(RRTMGSW_ERR_ARGUMENT, "cloud inputs are required when icld >= 1");
  if (icld >= 4 && !alpha) return fail(RRTMGSW_ERR_ARGUMENT, "alpha is required when icld = 4 or 5");
  if (iaer == 10 && (!tauaer || !ssaaer || !asmaer))
    return fail(RRTMGSW_ERR_ARGUMENT, "tauaer, ssaaer and asmaer are required when iaer = 10");
  if (iaer == 6 && !ecaer) return fail(RRTMGSW_ERR_ARGUMENT, "ecaer is required when iaer = 6");
  const uint32_t *mt = nullptr, *jumps = nullptr;
  if (icld >= 1 && prepare)
    if (int rc = mcica_prepare(ncol, nlay, icld, permuteseed, ir, s, &mt, &jumps)) return rc;
  a.mt_stream = mt;
  t.kiss_jump = jumps;
  return 0;
}

int rrtmgsw_mcica_sw_dev(int ncol, int nlay, int icld, int iaer, int permuteseed, int *irng,
                         const double *play, const double *plev, const double *tlay,
                         const double *tlev, const double *tsfc, const double *h2ovmr,
                         const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                         const double *n2ovmr, const double *o2vmr, const double *asdir,
                         const double *asdif, const double *aldir, const double *aldif,
                         const double *coszen, double adjes, int dyofyr, double scon,
                         int isolvar, int inflgsw, int iceflgsw, int liqflgsw,
                         const double *cldfrac, const double *ciwp, const double *clwp,
                         const double *rei, const double *rel, const double *tauc,
                         const double *ssac, const double *asmc, const double *fsfc,
                         const double *alpha, const double *tauaer, const double *ssaaer,
                         const double *asmaer, const double *ecaer, double *swuflx,
                         double *swdflx, double *swhr, double *swuflxc, double *swdflxc,
                         double *swhrc, const double *bndsolvar, const double *indsolvar,
                         const double *solcycfrac, int precision, void *stream) {
  if (int rc = check_init()) return rc;
  if (precision != RRTMGSW_FP64 && precision != RRTMGSW_MIXED && precision != RRTMGSW_FP64_EXP)
    return fail(RRTMGSW_ERR_ARGUMENT, "precision must be RRTMGSW_FP64, RRTMGSW_MIXED or RRTMGSW_FP64_EXP");
  hipStream_t s = (hipStream_t)stream;
  auto call = stream_call(s);
  RtArgs a{};
  DeviceTables t = ctx.dev;
  if (ncol > 0) {
    const void *outs[] = {swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc};
    for (const void *p : outs)
      if (!p) return fail(RRTMGSW_ERR_ARGUMENT, "a required array argument is NULL");
  }
  if (int rc = fused_setup(ncol, nlay, icld, iaer, permuteseed, irng, play, plev, tlay, h2ovmr, o3vmr,
                           co2vmr, ch4vmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr,
                           scon, isolvar, inflgsw, iceflgsw, liqflgsw, cldfrac, ciwp, clwp, rei, rel,
                           tauc, ssac, asmc, fsfc, alpha, tauaer, ssaaer, asmaer, ecaer, bndsolvar,
                           indsolvar, solcycfrac, s, a, t))
    return rc;
  (void)tlev, (void)tsfc, (void)n2ovmr;
  if (ncol == 0) return 0;
  a.swuflx = swuflx, a.swdflx = swdflx, a.swhr = swhr, a.swuflxc = swuflxc, a.swdflxc = swdflxc;
  a.swhrc = swhrc;
  return run_rt(a, t, s, precision);
}

int rrtmgsw_setcoef_dev(int ncol, int nlay, const double *play, const double *plev, const double *tlay,
                        const double *h2ovmr, const double *o3vmr, const double *co2vmr,
                        const double *ch4vmr, const double *o2vmr, int *laytrop, int *laylow,
                        int *jp, int *jt, int *jt1, double *colch4, double *colco2, double *colh2o,
                        double *colmol, double *colo2, double *colo3, double *fac00, double *fac01,
                        double *fac10, double *fac11, double *selffac, double *selffrac,
                        int *indself, double *forfac, double *forfrac, int *indfor, int *laysolfr,
                        void *stream) {
  if (int rc = check_init()) return rc;
  if (ncol < 0 || nlay < 1 || nlay > MAX_LAY)
    return fail(RRTMGSW_ERR_ARGUMENT, "nlay must lie in 1..256 and ncol >= 0");
  if (ncol == 0) return 0;
  const void *need[] = {play, plev, tlay, h2ovmr, o3vmr, co2vmr, ch4vmr, o2vmr};
  for (const void *p : need)
    if (!p) return fail(RRTMGSW_ERR_ARGUMENT, "a required array argument is NULL");
  hipStream_t s = (hipStream_t)stream;
  RtArgs a{};
  a.ncol = ncol, a.nlay = nlay, a.col0 = 0, a.ncol_chunk = ncol;
  a.play = play, a.plev = plev, a.tlay = tlay, a.h2ovmr = h2ovmr, a.o3vmr = o3vmr;
  a.co2vmr = co2vmr, a.ch4vmr = ch4vmr, a.o2vmr = o2vmr;
  SetcoefOut o{};
  o.laytrop = laytrop, o.laylow = laylow, o.laysolfr = laysolfr;
  o.jp = jp, o.jt = jt, o.jt1 = jt1, o.indself = indself, o.indfor = indfor;
  o.fac00 = fac00, o.fac01 = fac01, o.fac10 = fac10, o.fac11 = fac11;
  o.selffac = selffac, o.selffrac = selffrac, o.forfac = forfac, o.forfrac = forfrac;
  o.colh2o = colh2o, o.colco2 = colco2, o.colo3 = colo3, o.colch4 = colch4, o.colo2 = colo2;
  o.colmol = colmol;
  HIPCHK(launch_setcoef(a, o, ctx.dev, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int rrtmgsw_mcica_sw_dev_noalloc(int ncol, int nlay, int icld, int iaer, int permuteseed, int *irng,
                                 const double *play, const double *plev, const double *tlay,
                                 const double *tlev, const double *tsfc, const double *h2ovmr,
                                 const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                                 const double *n2ovmr, const double *o2vmr, const double *asdir,
                                 const double *asdif, const double *aldir, const double *aldif,
                                 const double *coszen, double adjes, int dyofyr, double scon,
                                 int isolvar, int inflgsw, int iceflgsw, int liqflgsw,
                                 const double *cldfrac, const double *ciwp, const double *clwp,
                                 const double *rei, const double *rel, const double *tauc,
                                 const double *ssac, const double *asmc, const double *fsfc,
                                 const double *alpha, const double *tauaer, const double *ssaaer,
                                 const double *asmaer, const double *ecaer, double *swuflx,
                                 double *swdflx, double *swhr, double *swuflxc, double *swdflxc,
                                 double *swhrc, const double *bndsolvar, const double *indsolvar,
                                 const double *solcycfrac, int precision, void *stream) {
  if (int rc = check_init()) return rc;
  if (precision != RRTMGSW_FP64 && precision != RRTMGSW_MIXED && precision != RRTMGSW_FP64_EXP)
    return fail(RRTMGSW_ERR_ARGUMENT, "precision must be RRTMGSW_FP64, RRTMGSW_MIXED or RRTMGSW_FP64_EXP");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  char *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(ctx.mu);
    auto it = ctx.noalloc.find(dev);
    if (it == ctx.noalloc.end())
      return fail(RRTMGSW_ERR_ARGUMENT, "the current device is not the one rrtmgsw_init set up");
    base = it->second;
  }
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> one(ctx.noalloc_mu);
  RtArgs a{};
  DeviceTables t = ctx.dev;
  if (ncol > 0) {
    const void *outs[] = {swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc};
    for (const void *p : outs)
      if (!p) return fail(RRTMGSW_ERR_ARGUMENT, "a required array argument is NULL");
  }
  if (int rc = fused_setup(ncol, nlay, icld, iaer, permuteseed, irng, play, plev, tlay, h2ovmr, o3vmr,
                           co2vmr, ch4vmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr,
                           scon, isolvar, inflgsw, iceflgsw, liqflgsw, cldfrac, ciwp, clwp, rei, rel,
                           tauc, ssac, asmc, fsfc, alpha, tauaer, ssaaer, asmaer, ecaer, bndsolvar,
                           indsolvar, solcycfrac, s, a, t, false))
    return rc;
  (void)tlev, (void)tsfc, (void)n2ovmr;
  if (ncol == 0) return 0;
  // the array: status word | KISS jumps (112 records) | MT19937 stream (irng = 1) | per-column
  // elimination coefficients and cloud optics of one launch's columns
  const size_t jump_off = 256, mt_off = jump_off + ((NGPT * sizeof(KissJump) + 255) / 256) * 256;
  const int per_sub = icld == 3 ? 1 : (icld >= 4 ? 2 * nlay : nlay);
  const bool mt = icld >= 1 && a.mc_irng == 1;
  const size_t ndraw = mt ? (size_t)NGPT * ncol * per_sub : 0;
  const size_t rt_off = mt_off + ((ndraw * 4 + 255) / 256) * 256;
  const size_t per_col = scr_alloc_doubles(nlay) * sizeof(double);
  if (rt_off + per_col > RRTMGSW_NOALLOC_SCRATCH_BYTES)
    return fail(RRTMGSW_ERR_ARGUMENT,
                "rrtmgsw_mcica_sw_dev_noalloc: the Mersenne Twister stream of this call (" +
                    std::to_string(ndraw * 4) + " bytes) does not fit its " +
                    std::to_string((size_t)RRTMGSW_NOALLOC_SCRATCH_BYTES) + "-byte scratch");
  const int chunk = (int)std::min<size_t>(
      (size_t)ncol, launch_column_limit((RRTMGSW_NOALLOC_SCRATCH_BYTES - rt_off) / per_col));
  if (icld >= 1) {
    if (a.mc_irng == 0) {
      // the per-sub-column jumps, formed on the GPU ahead of the columns' launches
      if (int rc = kiss_jumps(permuteseed, per_sub, s, reinterpret_cast<KissJump *>(base + jump_off)))
        return rc;
      t.kiss_jump = reinterpret_cast<const uint32_t *>(base + jump_off);
    } else {
      // one workgroup generates the stream in order (no jump-ahead table)
      uint32_t *w = reinterpret_cast<uint32_t *>(base + mt_off);
      HIPCHK(launch_mt_stream((uint32_t)permuteseed, (long long)ndraw, w, nullptr, s));
      a.mt_stream = w;
    }
  }
  a.swuflx = swuflx, a.swdflx = swdflx, a.swhr = swhr, a.swuflxc = swuflxc, a.swdflxc = swdflxc;
  a.swhrc = swhrc;
  a.status = reinterpret_cast<int *>(base);
  a.scratch = reinterpret_cast<double *>(base + rt_off);
  hipError_t e = launch_rt_chunks(a, t, s, precision, chunk);
  if (e != hipSuccess) return fail(RRTMGSW_ERR_HIP, hipGetErrorString(e));
  return finish(s, a.status);
}

size_t rrtmgsw_adjoint_work_bytes(int nlay, int ckpt) {
  if (nlay < 1 || ckpt < 0) return 0;
  return adj_work_bytes(nlay, ckpt);
}

int rrtmgsw_mcica_sw_adjoint_dev(int ncol, int nlay, int icld, int iaer, int permuteseed, int *irng,
                                 const double *play, const double *plev, const double *tlay,
                                 const double *tlev, const double *tsfc, const double *h2ovmr,
                                 const double *o3vmr, const double *co2vmr, const double *ch4vmr,
                                 const double *n2ovmr, const double *o2vmr, const double *asdir,
                                 const double *asdif, const double *aldir, const double *aldif,
                                 const double *coszen, double adjes, int dyofyr, double scon,
                                 int isolvar, int inflgsw, int iceflgsw, int liqflgsw,
                                 const double *cldfrac, const double *ciwp, const double *clwp,
                                 const double *rei, const double *rel, const double *tauc,
                                 const double *ssac, const double *asmc, const double *fsfc,
                                 const double *alpha, const double *tauaer, const double *ssaaer,
                                 const double *asmaer, const double *ecaer,
                                 const double *bndsolvar, const double *indsolvar,
                                 const double *solcycfrac, const double *seed_toa,
                                 const double *seed_hr, double *d_tlay, double *d_h2ovmr,
                                 double *d_o3vmr, double *d_clwp, double *d_ciwp, double *d_alb,
                                 int ckpt, void *stream) {
  if (int rc = check_init()) return rc;
  hipStream_t s = (hipStream_t)stream;
  auto call = stream_call(s);
  RtArgs a{};
  DeviceTables t = ctx.dev;
  if (ncol > 0) {
    const void *need[] = {seed_toa, seed_hr, d_tlay, d_h2ovmr, d_o3vmr, d_clwp, d_ciwp, d_alb};
    for (const void *p : need)
      if (!p) return fail(RRTMGSW_ERR_ARGUMENT, "a required array argument is NULL");
  }
  if (int rc = fused_setup(ncol, nlay, icld, iaer, permuteseed, irng, play, plev, tlay, h2ovmr, o3vmr,
                           co2vmr, ch4vmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr,
                           scon, isolvar, inflgsw, iceflgsw, liqflgsw, cldfrac, ciwp, clwp, rei, rel,
                           tauc, ssac, asmc, fsfc, alpha, tauaer, ssaaer, asmaer, ecaer, bndsolvar,
                           indsolvar, solcycfrac, s, a, t))
    return rc;
  (void)tlev, (void)tsfc, (void)n2ovmr;
  if (ncol == 0) return 0;
  AdjArgs A{};
  A.seed_toa = seed_toa, A.seed_hr = seed_hr;
  A.d_tlay = d_tlay, A.d_h2ovmr = d_h2ovmr, A.d_o3vmr = d_o3vmr;
  A.d_clwp = d_clwp, A.d_ciwp = d_ciwp, A.d_alb = d_alb;
  if (ckpt < 0 || ckpt > nlay) return fail(RRTMGSW_ERR_ARGUMENT, "ckpt must lie in 0..nlay");
  A.ckpt = ckpt;
  {
    // an adjoint workgroup needs more than 64 KB of LDS from 225 layers on (74 KB at 256)
    int dev = 0, lds_max = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (adj_lds_bytes(nlay) > (size_t)lds_max)
      return fail(RRTMGSW_ERR_ARGUMENT, "nlay needs more LDS per workgroup than this device has");
  }
  const size_t per_col = adj_work_bytes(nlay, ckpt);
  int chunk = (int)std::min<size_t>((size_t)ncol, launch_column_limit(std::max<size_t>(64, ((size_t)4 << 30) / per_col)));
  void *work = nullptr, *st = nullptr;
  {
    std::lock_guard<std::mutex> lk(ctx.mu);
    if (int rc = stream_buffer(ctx.adjwork, s, (size_t)chunk * per_col, &work)) return rc;
    if (int rc = stream_buffer(ctx.status, s, 256, &st)) return rc;
  }
  hipLaunchKernelGGL(zero_int, dim3(1), dim3(1), 0, s, (int *)st);
  a.status = (int *)st;
  A.work = (double *)work;
  for (int c0 = 0; c0 < ncol; c0 += chunk) {
    a.col0 = c0;
    a.ncol_chunk = std::min(chunk, ncol - c0);
    A.f = a;
    HIPCHK(launch_adjoint(A, t, s));
  }
  return finish(s, (int *)st);
}

}  // extern "C"
