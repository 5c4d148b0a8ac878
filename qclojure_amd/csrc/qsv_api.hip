// qsv_api.hip — the C ABI of libqsv (include/qsv.h).
//
// Decodes qsv_op records exactly as the reference dispatcher
// apply-gate-to-state does (src/org/soulspace/qclojure/domain/circuit.clj:933-1072)
// and turns each into one streaming kernel on the state's own HIP stream.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "qsv_internal.h"

using namespace qsv;

namespace {
thread_local std::string g_err;
std::mutex g_mu;
bool g_init = false;
int g_device = 0;

int fail(int code, const std::string& msg)
{
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return fail(QSV_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Pending {
  int cls;
  hipEvent_t a, b;
  double bytes;
};
}  // namespace

struct qsv_state {
  int n = 0;
  int prec = QSV_C128;
  int device = 0;
  uint64_t S = 0;
  void* amps = nullptr;
  hipStream_t stream = nullptr;
  double* partials = nullptr;   // kReduceBlocks * 2 * kMaxPauliPerPass
  double* dout = nullptr;       // 64 doubles of reduction results
  double* prefix = nullptr;     // sampling chunk prefix (lazy)
  double* hist = nullptr;       // measurement histogram (lazy)
  uint64_t hist_cap = 0;
  bool timing = false;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> evpool;
  qsv_kernel_stat stats[QSV_KCLASS_COUNT] = {};
  uint64_t bytes_per_amp() const { return prec == QSV_C128 ? 16 : 8; }
};

namespace {

int ensure_init()
{
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_init) return QSV_OK;
  int count = 0;
  HIPCHK(hipGetDeviceCount(&count));
  if (count <= 0) return fail(QSV_E_DEVICE, "no HIP device visible");
  HIPCHK(hipSetDevice(g_device));
  g_init = true;
  return QSV_OK;
}

hipEvent_t take_event(qsv_state* s)
{
  if (!s->evpool.empty()) {
    hipEvent_t e = s->evpool.back();
    s->evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int flush_timing(qsv_state* s)
{
  if (s->pending.empty()) return QSV_OK;
  HIPCHK(hipEventSynchronize(s->pending.back().b));
  for (auto& p : s->pending) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
    s->stats[p.cls].launches += 1;
    s->stats[p.cls].total_ms += ms;
    s->stats[p.cls].bytes += p.bytes;
    s->evpool.push_back(p.a);
    s->evpool.push_back(p.b);
  }
  s->pending.clear();
  return QSV_OK;
}

template <class F>
int timed(qsv_state* s, int cls, double bytes, F&& launch)
{
  if (!s->timing) {
    launch();
    return QSV_OK;
  }
  hipEvent_t a = take_event(s), b = take_event(s);
  if (!a || !b) return fail(QSV_E_DEVICE, "hipEventCreate failed");
  HIPCHK(hipEventRecord(a, s->stream));
  launch();
  HIPCHK(hipEventRecord(b, s->stream));
  s->pending.push_back({cls, a, b, bytes});
  if (s->pending.size() >= 2048) return flush_timing(s);
  return QSV_OK;
}

InsArgs make_ins(uint64_t S, std::vector<int> bits, uint64_t orMask)
{
  std::sort(bits.begin(), bits.end());
  InsArgs a{};
  a.nins = (int)bits.size();
  for (size_t k = 0; k < bits.size(); ++k) a.ins[k] = bits[k];
  a.count = S >> bits.size();
  a.orMask = orMask;
  return a;
}

void gate_matrix(int kind, double angle, double out[8])
{
  // gate.clj:38-283, row-major (re, im).
  double m[8] = {1, 0, 0, 0, 0, 0, 1, 0};
  const double s2 = 1.0 / std::sqrt(2.0);
  switch (kind) {
  case QSV_OP_X: m[0] = 0; m[2] = 1; m[4] = 1; m[6] = 0; break;
  case QSV_OP_Y: m[0] = 0; m[3] = -1; m[5] = 1; m[6] = 0; break;
  case QSV_OP_Z: m[6] = -1; break;
  case QSV_OP_H: m[0] = s2; m[2] = s2; m[4] = s2; m[6] = -s2; break;
  case QSV_OP_S: m[6] = 0; m[7] = 1; break;
  case QSV_OP_SDG: m[6] = 0; m[7] = -1; break;
  case QSV_OP_T: m[6] = std::cos(M_PI / 4); m[7] = std::sin(M_PI / 4); break;
  case QSV_OP_TDG: m[6] = std::cos(M_PI / -4); m[7] = std::sin(M_PI / -4); break;
  case QSV_OP_PHASE: m[6] = std::cos(angle); m[7] = std::sin(angle); break;
  case QSV_OP_RX: {
    const double c = std::cos(angle / 2), s = std::sin(angle / 2);
    m[0] = c; m[3] = -s; m[5] = -s; m[6] = c; break;
  }
  case QSV_OP_RY: {
    const double c = std::cos(angle / 2), s = std::sin(angle / 2);
    m[0] = c; m[2] = -s; m[4] = s; m[6] = c; break;
  }
  case QSV_OP_RZ:
    m[0] = std::cos(angle / -2); m[1] = std::sin(angle / -2);
    m[6] = std::cos(angle / 2); m[7] = std::sin(angle / 2); break;
  default: break;
  }
  std::memcpy(out, m, sizeof(m));
}

inline bool is_diag(const double m[8]) { return m[2] == 0 && m[3] == 0 && m[4] == 0 && m[5] == 0; }
inline bool is_one(double re, double im) { return re == 1.0 && im == 0.0; }

// Bits below this stay inside the wavefront (their 2^thr-amplitude runs are
// shorter than a 128 B line); bits at or above it are enumerated by insertion.
inline int low_bit_threshold(const qsv_state* s) { return s->prec == QSV_C128 ? 3 : 4; }

// 2x2 `m` applied to target bit tb where all control bits are set (no controls
// for a plain 1-qubit gate). Diagonal matrices take the phase kernel, which
// touches only the |1> half of the target when m00 == 1.
// Algorithmic bytes: each touched amplitude is read once and written once.
int apply_ctl_bits(qsv_state* s, const std::vector<int>& cbits, int tb, const double m[8])
{
  const uint64_t B = s->bytes_per_amp();
  const int thr = low_bit_threshold(s);
  std::vector<int> ins;
  uint64_t hor = 0, lpred = 0;
  for (int c : cbits) {
    if (c >= thr) { ins.push_back(c); hor |= 1ull << c; }
    else lpred |= 1ull << c;
  }
  double touched = (double)(s->S >> cbits.size());
  GateArgs g{};
  if (is_diag(m)) {
    const double d0[2] = {m[0], m[1]}, d1[2] = {m[6], m[7]};
    if (is_one(m[0], m[1])) {
      touched /= 2;
      if (tb >= thr) { ins.push_back(tb); hor |= 1ull << tb; }
      else lpred |= 1ull << tb;
    }
    g.ins = make_ins(s->S, ins, hor);
    g.lo_pred = lpred;
    g.tbit = 1ull << tb;
    return timed(s, QSV_KCLASS_DIAG, 2.0 * touched * B,
                 [&] { launch_phase(s->prec, s->amps, g, d0, d1, s->stream); });
  }
  if (tb >= thr) { ins.push_back(tb); g.tbit = 1ull << tb; }
  else g.lo_t = 1ull << tb;
  g.ins = make_ins(s->S, ins, hor);
  g.lo_pred = lpred;
  return timed(s, QSV_KCLASS_PAIR, 2.0 * touched * B,
               [&] { launch_gate1(s->prec, s->amps, g, m, s->stream); });
}

int apply_1q_bit(qsv_state* s, int b, const double m[8]) { return apply_ctl_bits(s, {}, b, m); }

// Exchange of index bits b1, b2 (times ph) where all control bits are set.
int apply_swap_bits(qsv_state* s, const std::vector<int>& cbits, int b1, int b2, const double ph[2])
{
  const uint64_t B = s->bytes_per_amp();
  const int thr = low_bit_threshold(s);
  std::vector<int> ins;
  uint64_t hor = 0, lpred = 0;
  for (int c : cbits) {
    if (c >= thr) { ins.push_back(c); hor |= 1ull << c; }
    else lpred |= 1ull << c;
  }
  SwapArgs a{};
  if (b1 < thr && b2 >= thr) std::swap(b1, b2);
  if (b1 >= thr && b2 >= thr) {
    ins.push_back(b1);
    ins.push_back(b2);
    a.hA = 1ull << b1;
    a.hB = 1ull << b2;
  } else if (b1 >= thr) {
    ins.push_back(b1);
    a.hA = 1ull << b1;
    a.lB = 1ull << b2;
  } else {
    a.lA = 1ull << b1;
    a.lB = 1ull << b2;
  }
  a.ins = make_ins(s->S, ins, hor);
  a.lo_pred = lpred;
  const double touched = (double)(s->S >> (cbits.size() + 1));
  return timed(s, QSV_KCLASS_PERM, 2.0 * touched * B,
               [&] { launch_swap2(s->prec, s->amps, a, ph, s->stream); });
}

// The reference's apply-controlled-gate applies the transpose of g
// (gate.clj:461-486): new[i0] = g00 a0 + g10 a1, new[i1] = g01 a0 + g11 a1.
void transpose2(const double g[8], double t[8])
{
  t[0] = g[0]; t[1] = g[1];
  t[2] = g[4]; t[3] = g[5];
  t[4] = g[2]; t[5] = g[3];
  t[6] = g[6]; t[7] = g[7];
}

int measure_impl(qsv_state* s, uint64_t mask, double u, uint64_t* outcome, double* probs_out)
{
  const int n = s->n;
  if (mask == 0 || (n < 64 && (mask >> n) != 0)) return fail(QSV_E_INVALID, "measure: qubit set empty or out of range");
  std::vector<int> pos;
  for (int q = 0; q < n; ++q)
    if ((mask >> q) & 1ull) pos.push_back(n - 1 - q);
  const int m = (int)pos.size();
  if (m > 20) return fail(QSV_E_UNSUPPORTED, "measure: more than 20 measured qubits");
  const uint64_t no = 1ull << m;
  if (s->hist_cap < no) {
    if (s->hist) hipFree(s->hist);
    s->hist = nullptr;
    HIPCHK(hipMalloc(&s->hist, no * sizeof(double)));
    s->hist_cap = no;
  }
  int rc = timed(s, QSV_KCLASS_OTHER, (double)s->S * s->bytes_per_amp(),
                 [&] { launch_outcome_hist(s->prec, s->amps, s->S, m, pos.data(), s->hist, s->stream); });
  if (rc) return rc;
  std::vector<double> pr(no);
  HIPCHK(hipMemcpyAsync(pr.data(), s->hist, no * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  // selection as measure-specific-qubits (state.clj:962-971)
  double total = 0.0;
  for (uint64_t o = 0; o < no; ++o) total += pr[o];
  const double r = u * total;
  double cum = 0.0;
  uint64_t sel = 0;
  for (uint64_t o = 0; o < no; ++o) {
    cum += pr[o];
    if (cum < r) sel = o + 1; else break;
  }
  if (sel > no - 1) sel = no - 1;
  const double p = pr[sel];
  const double f = p > 0 ? 1.0 / std::sqrt(p) : 1.0;
  rc = timed(s, QSV_KCLASS_OTHER, 2.0 * s->S * s->bytes_per_amp(),
             [&] { launch_collapse(s->prec, s->amps, s->S, m, pos.data(), sel, f, s->stream); });
  if (rc) return rc;
  if (outcome) *outcome = sel;
  if (probs_out) std::memcpy(probs_out, pr.data(), no * sizeof(double));
  return QSV_OK;
}

bool qok(int q, int n) { return q >= 0 && q < n; }

int apply_op(qsv_state* s, const qsv_op& op, uint32_t flags, uint64_t seed, uint64_t* measure_k)
{
  const int n = s->n;
  const int k = op.kind;
  double m[8], t[8];
  auto bit = [n](int q) { return n - 1 - q; };
  switch (k) {
  case QSV_OP_X: case QSV_OP_Y: case QSV_OP_Z: case QSV_OP_H: case QSV_OP_S: case QSV_OP_SDG:
  case QSV_OP_T: case QSV_OP_TDG: case QSV_OP_RX: case QSV_OP_RY: case QSV_OP_RZ: case QSV_OP_PHASE:
    if (!qok(op.q[0], n)) return fail(QSV_E_INVALID, "target qubit out of range");
    gate_matrix(k, op.angle, m);
    return apply_1q_bit(s, bit(op.q[0]), m);
  case QSV_OP_U:
    if (!qok(op.q[0], n)) return fail(QSV_E_INVALID, "target qubit out of range");
    return apply_1q_bit(s, bit(op.q[0]), op.m);
  case QSV_OP_CNOT: case QSV_OP_CZ: case QSV_OP_CY: case QSV_OP_RYDBERG_CZ:
  case QSV_OP_CRX: case QSV_OP_CRY: case QSV_OP_CRZ: {
    const int c = op.q[0], tq = op.q[1];
    if (!qok(c, n) || !qok(tq, n) || c == tq) return fail(QSV_E_INVALID, "control/target invalid");
    const int gk = (k == QSV_OP_CNOT) ? QSV_OP_X
                 : (k == QSV_OP_CZ || k == QSV_OP_RYDBERG_CZ) ? QSV_OP_Z
                 : (k == QSV_OP_CY) ? QSV_OP_Y
                 : (k == QSV_OP_CRX) ? QSV_OP_RX
                 : (k == QSV_OP_CRY) ? QSV_OP_RY : QSV_OP_RZ;
    gate_matrix(gk, op.angle, m);
    transpose2(m, t);
    return apply_ctl_bits(s, {bit(c)}, bit(tq), t);
  }
  case QSV_OP_CU: {
    const int c = op.q[0], tq = op.q[1];
    if (!qok(c, n) || !qok(tq, n) || c == tq) return fail(QSV_E_INVALID, "control/target invalid");
    return apply_ctl_bits(s, {bit(c)}, bit(tq), op.m);
  }
  case QSV_OP_SWAP: case QSV_OP_ISWAP: {
    const int q1 = op.q[0], q2 = op.q[1];
    if (!qok(q1, n) || !qok(q2, n) || q1 == q2) return fail(QSV_E_INVALID, "swap qubits invalid");
    const bool msb = (flags & QSV_FLAG_MSB_SWAP) != 0;
    const int b1 = msb ? bit(q1) : q1, b2 = msb ? bit(q2) : q2;
    const double ph[2] = {k == QSV_OP_ISWAP ? 0.0 : 1.0, k == QSV_OP_ISWAP ? 1.0 : 0.0};
    return apply_swap_bits(s, {}, b1, b2, ph);
  }
  case QSV_OP_TOFFOLI: {
    const int a = op.q[0], b = op.q[1], c = op.q[2];
    if (!qok(a, n) || !qok(b, n) || !qok(c, n) || a == b || a == c || b == c)
      return fail(QSV_E_INVALID, "toffoli qubits invalid");
    gate_matrix(QSV_OP_X, 0, m);
    return apply_ctl_bits(s, {bit(a), bit(b)}, bit(c), m);
  }
  case QSV_OP_FREDKIN: {
    const int a = op.q[0], b = op.q[1], c = op.q[2];
    if (!qok(a, n) || !qok(b, n) || !qok(c, n) || a == b || a == c || b == c)
      return fail(QSV_E_INVALID, "fredkin qubits invalid");
    const double ph[2] = {1.0, 0.0};
    return apply_swap_bits(s, {bit(a)}, bit(b), bit(c), ph);
  }
  case QSV_OP_RYDBERG_CPHASE: {
    const int c = op.q[0], tq = op.q[1];
    if (!qok(c, n) || !qok(tq, n) || c == tq) return fail(QSV_E_INVALID, "control/target invalid");
    const double dm[8] = {1, 0, 0, 0, 0, 0, std::cos(op.angle), std::sin(op.angle)};
    return apply_ctl_bits(s, {bit(c)}, bit(tq), dm);
  }
  case QSV_OP_RYDBERG_BLOCKADE: {
    if (op.mask == 0 || (n < 64 && (op.mask >> n) != 0)) return fail(QSV_E_INVALID, "blockade qubit set invalid");
    uint64_t bm = 0;
    for (int q = 0; q < n; ++q)
      if ((op.mask >> q) & 1ull) bm |= 1ull << bit(q);
    const double ph[2] = {std::cos(op.angle), std::sin(op.angle)};
    return timed(s, QSV_KCLASS_DIAG, 2.0 * s->S * s->bytes_per_amp(),
                 [&] { launch_blockade(s->prec, s->amps, s->S, bm, ph, s->stream); });
  }
  case QSV_OP_GLOBAL_RX: case QSV_OP_GLOBAL_RY: case QSV_OP_GLOBAL_RZ: case QSV_OP_GLOBAL_H:
  case QSV_OP_GLOBAL_X: case QSV_OP_GLOBAL_Y: case QSV_OP_GLOBAL_Z: {
    int gk;
    double ang = op.angle;
    switch (k) {
    case QSV_OP_GLOBAL_RX: gk = QSV_OP_RX; break;
    case QSV_OP_GLOBAL_RY: gk = QSV_OP_RY; break;
    case QSV_OP_GLOBAL_RZ: gk = QSV_OP_RZ; break;
    case QSV_OP_GLOBAL_H: gk = QSV_OP_H; break;
    case QSV_OP_GLOBAL_X: gk = QSV_OP_RX; ang = M_PI; break;
    case QSV_OP_GLOBAL_Y: gk = QSV_OP_RY; ang = M_PI; break;
    default: gk = QSV_OP_RZ; ang = M_PI; break;
    }
    gate_matrix(gk, ang, m);
    for (int q = 0; q < n; ++q) {
      int rc = apply_1q_bit(s, bit(q), m);
      if (rc) return rc;
    }
    return QSV_OK;
  }
  case QSV_OP_MEASURE: {
    const double u = qsv_uniform(seed, (*measure_k)++);
    return measure_impl(s, op.mask, u, nullptr, nullptr);
  }
  default:
    return fail(QSV_E_UNSUPPORTED, "Unknown gate type (kind " + std::to_string(k) + ")");
  }
}

int check_state(const qsv_state* s)
{
  if (!s || !s->amps) return fail(QSV_E_INVALID, "null or destroyed state");
  HIPCHK(hipSetDevice(s->device));
  return QSV_OK;
}

constexpr uint64_t kXferChunk = 1ull << 24;  // amplitudes per staged transfer

}  // namespace

extern "C" {

int qsv_init(int device)
{
  std::lock_guard<std::mutex> lk(g_mu);
  int count = 0;
  HIPCHK(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(QSV_E_INVALID, "device index out of range");
  HIPCHK(hipSetDevice(device));
  g_device = device;
  g_init = true;
  return QSV_OK;
}

int qsv_shutdown(void)
{
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_init) hipDeviceSynchronize();
  g_init = false;
  return QSV_OK;
}

int qsv_device_query(qsv_device_info* out)
{
  if (!out) return fail(QSV_E_INVALID, "null output");
  int rc = ensure_init();
  if (rc) return rc;
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, g_device));
  std::memset(out, 0, sizeof(*out));
  std::snprintf(out->name, sizeof(out->name), "%s", p.name);
  std::snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
  out->compute_units = p.multiProcessorCount;
  out->wavefront_size = p.warpSize;
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  out->hbm_bytes = tot;
  out->hbm_free_bytes = fr;
  return QSV_OK;
}

const char* qsv_last_error(void) { return g_err.c_str(); }

const char* qsv_version(void) { return "qsv 0.1.0 (gfx950)"; }

double qsv_uniform(uint64_t seed, uint64_t k)
{
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

int qsv_state_create(int num_qubits, int precision, qsv_state** out)
{
  if (!out) return fail(QSV_E_INVALID, "null output");
  *out = nullptr;
  if (num_qubits < 1 || num_qubits > 36) return fail(QSV_E_INVALID, "num_qubits must be in [1, 36]");
  if (precision != QSV_C128 && precision != QSV_C64) return fail(QSV_E_INVALID, "unknown precision");
  int rc = ensure_init();
  if (rc) return rc;
  auto* s = new qsv_state();
  s->n = num_qubits;
  s->prec = precision;
  s->device = g_device;
  s->S = 1ull << num_qubits;
  auto cleanup = [&](int code, const std::string& msg) {
    if (s->amps) hipFree(s->amps);
    if (s->partials) hipFree(s->partials);
    if (s->dout) hipFree(s->dout);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
    return fail(code, msg);
  };
  if (hipSetDevice(s->device) != hipSuccess) return cleanup(QSV_E_DEVICE, "hipSetDevice failed");
  if (hipMalloc(&s->amps, s->S * s->bytes_per_amp()) != hipSuccess)
    return cleanup(QSV_E_NOMEM, "hipMalloc of state vector failed (" + std::to_string(s->S * s->bytes_per_amp()) + " B)");
  if (hipMalloc(&s->partials, sizeof(double) * kReduceBlocks * 2 * kMaxPauliPerPass) != hipSuccess)
    return cleanup(QSV_E_NOMEM, "hipMalloc of reduction scratch failed");
  if (hipMalloc(&s->dout, sizeof(double) * 64) != hipSuccess)
    return cleanup(QSV_E_NOMEM, "hipMalloc of reduction output failed");
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(QSV_E_DEVICE, "hipStreamCreate failed");
  launch_fill_basis(s->prec, s->amps, s->S, 0, s->stream);
  if (hipStreamSynchronize(s->stream) != hipSuccess) return cleanup(QSV_E_DEVICE, "initialising |0..0> failed");
  *out = s;
  return QSV_OK;
}

int qsv_state_destroy(qsv_state* s)
{
  if (!s) return QSV_OK;
  hipSetDevice(s->device);
  if (s->stream) hipStreamSynchronize(s->stream);
  for (auto& p : s->pending) {
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  for (auto e : s->evpool) hipEventDestroy(e);
  if (s->amps) hipFree(s->amps);
  if (s->partials) hipFree(s->partials);
  if (s->dout) hipFree(s->dout);
  if (s->prefix) hipFree(s->prefix);
  if (s->hist) hipFree(s->hist);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
  return QSV_OK;
}

int qsv_state_num_qubits(const qsv_state* s) { return s ? s->n : QSV_E_INVALID; }
int qsv_state_precision(const qsv_state* s) { return s ? s->prec : QSV_E_INVALID; }

int qsv_state_init_basis(qsv_state* s, uint64_t index)
{
  int rc = check_state(s);
  if (rc) return rc;
  if (index >= s->S) return fail(QSV_E_INVALID, "basis index out of range");
  launch_fill_basis(s->prec, s->amps, s->S, index, s->stream);
  HIPCHK(hipGetLastError());
  return QSV_OK;
}

int qsv_state_upload(qsv_state* s, const double* amps, uint64_t offset, uint64_t count)
{
  int rc = check_state(s);
  if (rc) return rc;
  if (!amps || offset > s->S || count > s->S - offset) return fail(QSV_E_INVALID, "upload range invalid");
  if (s->prec == QSV_C128) {
    HIPCHK(hipMemcpyAsync((char*)s->amps + offset * 16, amps, count * 16, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return QSV_OK;
  }
  double* stage = nullptr;
  const uint64_t chunk = std::min<uint64_t>(count, kXferChunk);
  HIPCHK(hipMalloc(&stage, chunk * 16));
  for (uint64_t done = 0; done < count; done += chunk) {
    const uint64_t c = std::min(chunk, count - done);
    hipError_t e = hipMemcpyAsync(stage, amps + 2 * done, c * 16, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) launch_f64_to_c64(stage, (char*)s->amps + (offset + done) * 8, c, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
      hipFree(stage);
      return fail(QSV_E_DEVICE, std::string("upload: ") + hipGetErrorString(e));
    }
  }
  hipFree(stage);
  return QSV_OK;
}

int qsv_state_download(const qsv_state* s, double* amps, uint64_t offset, uint64_t count)
{
  int rc = check_state(s);
  if (rc) return rc;
  if (!amps || offset > s->S || count > s->S - offset) return fail(QSV_E_INVALID, "download range invalid");
  if (s->prec == QSV_C128) {
    HIPCHK(hipMemcpyAsync(amps, (const char*)s->amps + offset * 16, count * 16, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return QSV_OK;
  }
  double* stage = nullptr;
  const uint64_t chunk = std::min<uint64_t>(count, kXferChunk);
  HIPCHK(hipMalloc(&stage, chunk * 16));
  for (uint64_t done = 0; done < count; done += chunk) {
    const uint64_t c = std::min(chunk, count - done);
    launch_c64_to_f64((const char*)s->amps + (offset + done) * 8, stage, c, s->stream);
    hipError_t e = hipMemcpyAsync(amps + 2 * done, stage, c * 16, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
      hipFree(stage);
      return fail(QSV_E_DEVICE, std::string("download: ") + hipGetErrorString(e));
    }
  }
  hipFree(stage);
  return QSV_OK;
}

int qsv_state_copy(qsv_state* dst, const qsv_state* src)
{
  int rc = check_state(dst);
  if (rc) return rc;
  rc = check_state(src);
  if (rc) return rc;
  if (dst->n != src->n || dst->prec != src->prec) return fail(QSV_E_INVALID, "copy: shape or precision mismatch");
  HIPCHK(hipStreamSynchronize(src->stream));
  HIPCHK(hipMemcpyAsync(dst->amps, src->amps, src->S * src->bytes_per_amp(), hipMemcpyDeviceToDevice, dst->stream));
  HIPCHK(hipStreamSynchronize(dst->stream));
  return QSV_OK;
}

int qsv_state_sync(qsv_state* s)
{
  int rc = check_state(s);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  return QSV_OK;
}

int qsv_apply_program(qsv_state* s, const qsv_op* ops, uint64_t n_ops, uint32_t flags, uint64_t seed)
{
  int rc = check_state(s);
  if (rc) return rc;
  if (n_ops && !ops) return fail(QSV_E_INVALID, "null ops");
  uint64_t mk = 0;
  for (uint64_t i = 0; i < n_ops; ++i) {
    rc = apply_op(s, ops[i], flags, seed, &mk);
    if (rc) {
      g_err = "op " + std::to_string(i) + ": " + g_err;
      return rc;
    }
  }
  HIPCHK(hipGetLastError());
  return QSV_OK;
}

int qsv_norm2(const qsv_state* s, double* out)
{
  int rc = check_state(s);
  if (rc) return rc;
  if (!out) return fail(QSV_E_INVALID, "null output");
  launch_norm2(s->prec, s->amps, s->S, s->partials, s->dout, s->stream);
  HIPCHK(hipMemcpyAsync(out, s->dout, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return QSV_OK;
}

int qsv_normalize(qsv_state* s)
{
  double n2 = 0;
  int rc = qsv_norm2(s, &n2);
  if (rc) return rc;
  const double norm = std::sqrt(n2);
  // normalize-state: divide only if norm > tolerance (state.clj:545-549)
  if (norm > 0 && norm > 1e-12) launch_scale(s->prec, s->amps, s->S, 1.0 / norm, s->stream);
  HIPCHK(hipGetLastError());
  return QSV_OK;
}

int qsv_probabilities(const qsv_state* s, uint64_t offset, uint64_t count, double* out)
{
  int rc = check_state(s);
  if (rc) return rc;
  if (!out || offset > s->S || count > s->S - offset) return fail(QSV_E_INVALID, "probability range invalid");
  if (count == 0) return QSV_OK;
  double* buf = nullptr;
  const uint64_t chunk = std::min<uint64_t>(count, kXferChunk);
  HIPCHK(hipMalloc(&buf, chunk * sizeof(double)));
  for (uint64_t done = 0; done < count; done += chunk) {
    const uint64_t c = std::min(chunk, count - done);
    launch_probs(s->prec, s->amps, offset + done, c, buf, s->stream);
    hipError_t e = hipMemcpyAsync(out + done, buf, c * sizeof(double), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
      hipFree(buf);
      return fail(QSV_E_DEVICE, std::string("probabilities: ") + hipGetErrorString(e));
    }
  }
  hipFree(buf);
  return QSV_OK;
}

int qsv_sample(const qsv_state* cs, uint64_t shots, uint64_t seed, uint64_t* outcomes)
{
  int rc = check_state(cs);
  if (rc) return rc;
  if (shots && !outcomes) return fail(QSV_E_INVALID, "null outcomes");
  auto* s = const_cast<qsv_state*>(cs);
  if (!s->prefix) {
    const uint64_t nch = std::max<uint64_t>(1, s->S >> kSampleChunkLog);
    HIPCHK(hipMalloc(&s->prefix, nch * sizeof(double)));
  }
  uint64_t* dout = nullptr;
  if (shots) HIPCHK(hipMalloc(&dout, shots * sizeof(uint64_t)));
  launch_sample(s->prec, s->amps, s->S, s->prefix, shots, seed, dout, s->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && shots)
    e = hipMemcpyAsync(outcomes, dout, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (dout) hipFree(dout);
  if (e != hipSuccess) return fail(QSV_E_DEVICE, std::string("sample: ") + hipGetErrorString(e));
  return QSV_OK;
}

int qsv_measure_qubits(qsv_state* s, uint64_t mask, double u, uint64_t* outcome, double* probs)
{
  int rc = check_state(s);
  if (rc) return rc;
  if (!(u >= 0.0 && u < 1.0)) return fail(QSV_E_INVALID, "uniform must lie in [0, 1)");
  rc = measure_impl(s, mask, u, outcome, probs);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  return QSV_OK;
}

int qsv_expect_pauli(const qsv_state* s, const char* strings, uint64_t n_terms, double* out)
{
  int rc = check_state(s);
  if (rc) return rc;
  if (n_terms && (!strings || !out)) return fail(QSV_E_INVALID, "null argument");
  const int n = s->n;
  std::vector<uint64_t> xm(n_terms), sm(n_terms);
  std::vector<int> ny(n_terms);
  for (uint64_t t = 0; t < n_terms; ++t) {
    const char* p = strings + t * (uint64_t)n;
    uint64_t x = 0, sg = 0;
    int y = 0;
    for (int q = 0; q < n; ++q) {
      const uint64_t b = 1ull << (n - 1 - q);
      switch (p[q]) {
      case 'I': break;
      case 'X': x |= b; break;
      case 'Y': x |= b; sg |= b; ++y; break;
      case 'Z': sg |= b; break;
      default: return fail(QSV_E_INVALID, "pauli string holds a character outside IXYZ");
      }
    }
    xm[t] = x; sm[t] = sg; ny[t] = y;
  }
  double res[2 * kMaxPauliPerPass];
  for (uint64_t t0 = 0; t0 < n_terms; t0 += kMaxPauliPerPass) {
    const int nt = (int)std::min<uint64_t>(kMaxPauliPerPass, n_terms - t0);
    launch_pauli(s->prec, s->amps, s->S, nt, &xm[t0], &sm[t0], s->partials, s->dout, s->stream);
    HIPCHK(hipMemcpyAsync(res, s->dout, sizeof(double) * 2 * nt, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int k = 0; k < nt; ++k) {
      // value = Re((-i)^nY * (re + i im))
      const double re = res[2 * k], im = res[2 * k + 1];
      double v;
      switch (ny[t0 + k] & 3) {
      case 0: v = re; break;
      case 1: v = im; break;
      case 2: v = -re; break;
      default: v = -im; break;
      }
      out[t0 + k] = v;
    }
  }
  return QSV_OK;
}

int qsv_expect_1q(const qsv_state* s, int qubit, const double m[8], double out[2])
{
  int rc = check_state(s);
  if (rc) return rc;
  if (!m || !out) return fail(QSV_E_INVALID, "null argument");
  if (!qok(qubit, s->n)) return fail(QSV_E_INVALID, "qubit out of range");
  launch_expect1q(s->prec, s->amps, s->S, 1ull << (s->n - 1 - qubit), m, s->partials, s->dout, s->stream);
  HIPCHK(hipMemcpyAsync(out, s->dout, sizeof(double) * 2, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return QSV_OK;
}

int qsv_inner(const qsv_state* a, const qsv_state* b, double out[2])
{
  int rc = check_state(a);
  if (rc) return rc;
  rc = check_state(b);
  if (rc) return rc;
  if (!out) return fail(QSV_E_INVALID, "null output");
  if (a->n != b->n || a->prec != b->prec) return fail(QSV_E_INVALID, "inner: shape or precision mismatch");
  HIPCHK(hipStreamSynchronize(b->stream));
  launch_inner(a->prec, a->amps, b->amps, a->S, a->partials, a->dout, a->stream);
  HIPCHK(hipMemcpyAsync(out, a->dout, sizeof(double) * 2, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  return QSV_OK;
}

int qsv_timing_enable(qsv_state* s, int enable)
{
  int rc = check_state(s);
  if (rc) return rc;
  if (!enable) {
    rc = flush_timing(s);
    if (rc) return rc;
  }
  s->timing = enable != 0;
  return QSV_OK;
}

int qsv_timing_read(qsv_state* s, qsv_kernel_stat out[QSV_KCLASS_COUNT], int reset)
{
  int rc = check_state(s);
  if (rc) return rc;
  rc = flush_timing(s);
  if (rc) return rc;
  if (out) std::memcpy(out, s->stats, sizeof(s->stats));
  if (reset) std::memset(s->stats, 0, sizeof(s->stats));
  return QSV_OK;
}

}  // extern "C"
