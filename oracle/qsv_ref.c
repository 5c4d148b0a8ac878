/*
 * qsv_ref.c — FP64 (complex128) CPU restatement of the reference ideal
 * simulator's gate-application and result-extraction path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the oracle: tests, the smoke check
 * and the CPU-baseline leg of the benchmark link it to check or time against.
 * The product library (libqsv) never links, loads or calls it.
 *
 * Two modes:
 *   literal (mode 0): the reference's own algorithms — every 1-qubit gate is
 *     expanded to the dense 2^n x 2^n Kronecker product I⊗..⊗G⊗..⊗I and applied
 *     by a dense mat-vec (gate.clj:323-395); controlled, permutation and
 *     Rydberg gates use the reference's accumulate-into-a-fresh-vector loops
 *     (gate.clj:424-486, 732-1095). Limited to n <= 12 (a 4096x4096 matrix).
 *   pair (mode 1): the same mathematics as in-place O(2^n) pair updates,
 *     OpenMP-parallel; used at sizes the literal mode cannot reach and as the
 *     multi-core CPU baseline.
 *
 * Conventions: qubit q is index bit n-1-q (state.clj:114-136); amplitudes are
 * interleaved (re, im) doubles.
 *
 * Pinned by: the reference's known-answer tests (gate_test.clj, state_test.clj,
 * hamiltonian_test.clj, ideal_simulator_test.clj, circuit_test.clj),
 * closed forms (GHZ, H^n, QFT-with-CRZ on basis states, QAOA p=1 edge formula),
 * and literal-vs-pair agreement on random circuits.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "../include/qsv.h"

typedef struct { double re, im; } rc;

static inline rc cmul(rc a, rc b) { rc r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re }; return r; }
static inline rc cadd(rc a, rc b) { rc r = { a.re + b.re, a.im + b.im }; return r; }
static inline rc mk(double re, double im) { rc r = { re, im }; return r; }

#define BIT(i, b) (((i) >> (b)) & 1ull)

/* ------------------------------------------------------------------------ */
/* Gate matrices, row-major (gate.clj:38-283).                               */
/* ------------------------------------------------------------------------ */
int ref_gate_matrix(int kind, double angle, double out[8])
{
  rc m[4] = { {1, 0}, {0, 0}, {0, 0}, {1, 0} };
  const double s2 = 1.0 / sqrt(2.0);                   /* hadamard, gate.clj:91-108 */
  switch (kind) {
  case QSV_OP_X: m[0] = mk(0, 0); m[1] = mk(1, 0); m[2] = mk(1, 0); m[3] = mk(0, 0); break;   /* :54 */
  case QSV_OP_Y: m[0] = mk(0, 0); m[1] = mk(0, -1); m[2] = mk(0, 1); m[3] = mk(0, 0); break;  /* :68 */
  case QSV_OP_Z: m[3] = mk(-1, 0); break;                                                     /* :82 */
  case QSV_OP_H: m[0] = mk(s2, 0); m[1] = mk(s2, 0); m[2] = mk(s2, 0); m[3] = mk(-s2, 0); break;
  case QSV_OP_S: m[3] = mk(0, 1); break;                                                      /* :150 */
  case QSV_OP_SDG: m[3] = mk(0, -1); break;                                                   /* :164 */
  case QSV_OP_T: m[3] = mk(cos(M_PI / 4), sin(M_PI / 4)); break;                              /* :177 */
  case QSV_OP_TDG: m[3] = mk(cos(M_PI / -4), sin(M_PI / -4)); break;                          /* :190 */
  case QSV_OP_PHASE: m[3] = mk(cos(angle), sin(angle)); break;                                /* :134-136 */
  case QSV_OP_RX: {                                                                           /* :215-221 */
    double c = cos(angle / 2), s = sin(angle / 2);
    m[0] = mk(c, 0); m[1] = mk(0, -s); m[2] = mk(0, -s); m[3] = mk(c, 0); break; }
  case QSV_OP_RY: {                                                                           /* :246-252 */
    double c = cos(angle / 2), s = sin(angle / 2);
    m[0] = mk(c, 0); m[1] = mk(-s, 0); m[2] = mk(s, 0); m[3] = mk(c, 0); break; }
  case QSV_OP_RZ:                                                                             /* :279-283 */
    m[0] = mk(cos(angle / -2), sin(angle / -2)); m[3] = mk(cos(angle / 2), sin(angle / 2)); break;
  default: return -1;
  }
  for (int k = 0; k < 4; ++k) { out[2 * k] = m[k].re; out[2 * k + 1] = m[k].im; }
  return 0;
}

static void load_m(const double in[8], rc m[4])
{
  for (int k = 0; k < 4; ++k) m[k] = mk(in[2 * k], in[2 * k + 1]);
}

/* ------------------------------------------------------------------------ */
/* Literal 1-qubit application: dense Kronecker expansion + mat-vec           */
/* (gate.clj:323-395; fastmath kronecker-product / matrix-vector-product).    */
/* ------------------------------------------------------------------------ */
#define LITERAL_MAX_QUBITS 12

static int literal_1q(rc* psi, int n, const rc g[4], int q)
{
  if (n > LITERAL_MAX_QUBITS) return -1;
  const uint64_t S = 1ull << n;
  /* reduce kronecker-product over the n factors, factor 0 outermost. */
  rc* K = (rc*)malloc(sizeof(rc) * S * S);
  rc* T = (rc*)malloc(sizeof(rc) * S * S);
  if (!K || !T) { free(K); free(T); return -2; }
  const rc I2[4] = { {1, 0}, {0, 0}, {0, 0}, {1, 0} };
  uint64_t dim = 2;
  const rc* f0 = (q == 0) ? g : I2;
  for (int k = 0; k < 4; ++k) K[k] = f0[k];
  for (int f = 1; f < n; ++f) {
    const rc* F = (f == q) ? g : I2;
    uint64_t nd = dim * 2;
    for (uint64_t i = 0; i < dim; ++i)
      for (uint64_t j = 0; j < dim; ++j)
        for (int k = 0; k < 2; ++k)
          for (int l = 0; l < 2; ++l)
            T[(i * 2 + k) * nd + (j * 2 + l)] = cmul(K[i * dim + j], F[k * 2 + l]);
    rc* t = K; K = T; T = t;
    dim = nd;
  }
  /* dense mat-vec, each row summed left to right. */
  rc* out = T;
  for (uint64_t i = 0; i < S; ++i) {
    rc acc = { 0, 0 };
    for (uint64_t j = 0; j < S; ++j) acc = cadd(acc, cmul(K[i * S + j], psi[j]));
    out[i] = acc;
  }
  memcpy(psi, out, sizeof(rc) * S);
  free(K); free(T);
  return 0;
}

/* Pair-update form of the same product: for every index pair differing in
 * bit b = n-1-q, (a0', a1') = G (a0, a1). */
static void pair_1q(rc* psi, int n, const rc g[4], int q)
{
  const int b = n - 1 - q;
  const uint64_t half = 1ull << (n - 1), tb = 1ull << b, lo = tb - 1;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < (int64_t)half; ++p) {
    uint64_t i0 = (((uint64_t)p & ~lo) << 1) | ((uint64_t)p & lo), i1 = i0 | tb;
    rc a0 = psi[i0], a1 = psi[i1];
    psi[i0] = cadd(cmul(g[0], a0), cmul(g[1], a1));
    psi[i1] = cadd(cmul(g[2], a0), cmul(g[3], a1));
  }
}

/* apply-controlled-gate (gate.clj:424-486), literal: accumulate into a fresh
 * zero vector. Note the reference adds a0*g01 into the flipped index, i.e.
 * the target pair is transformed by the TRANSPOSE of g. */
static int literal_controlled(rc* psi, int n, int c, int t, const rc g[4])
{
  const uint64_t S = 1ull << n;
  rc* r = (rc*)calloc(S, sizeof(rc));
  if (!r) return -2;
  const int cb = n - 1 - c, tb = n - 1 - t;
  for (uint64_t i = 0; i < S; ++i) {
    rc a = psi[i];
    if (BIT(i, cb)) {
      uint64_t fi = i ^ (1ull << tb);
      if (BIT(i, tb)) {
        r[fi] = cadd(r[fi], cmul(a, g[2]));
        r[i] = cadd(r[i], cmul(a, g[3]));
      } else {
        r[i] = cadd(r[i], cmul(a, g[0]));
        r[fi] = cadd(r[fi], cmul(a, g[1]));
      }
    } else {
      r[i] = cadd(r[i], a);
    }
  }
  memcpy(psi, r, sizeof(rc) * S);
  free(r);
  return 0;
}

/* In-place form of the same update: on control=1 pairs,
 * a0' = g00 a0 + g10 a1, a1' = g01 a0 + g11 a1. */
static void pair_controlled(rc* psi, int n, int c, int t, const rc g[4])
{
  const int cb = n - 1 - c, tb = n - 1 - t;
  const uint64_t S = 1ull << n;
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)S; ++ii) {
    uint64_t i = (uint64_t)ii;
    if (!BIT(i, cb) || BIT(i, tb)) continue;
    uint64_t j = i | (1ull << tb);
    rc a0 = psi[i], a1 = psi[j];
    psi[i] = cadd(cmul(a0, g[0]), cmul(a1, g[2]));
    psi[j] = cadd(cmul(a0, g[1]), cmul(a1, g[3]));
  }
}

/* swap-gate / iswap-gate (gate.clj:732-834): bit positions are the qubit
 * numbers themselves (LSB order) in the reference; msb=1 uses n-1-q. */
static int ref_swap_like(rc* psi, int n, int q1, int q2, int iswap, int msb)
{
  const uint64_t S = 1ull << n;
  const int b1 = msb ? n - 1 - q1 : q1, b2 = msb ? n - 1 - q2 : q2;
  rc* r = (rc*)malloc(sizeof(rc) * S);
  if (!r) return -2;
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)S; ++ii) {
    uint64_t i = (uint64_t)ii;
    uint64_t x1 = BIT(i, b1), x2 = BIT(i, b2);
    if (x1 == x2) { r[i] = psi[i]; continue; }
    uint64_t j = (i & ~(1ull << b1) & ~(1ull << b2)) | (x2 << b1) | (x1 << b2);
    r[i] = iswap ? cmul(mk(0, 1), psi[j]) : psi[j];
  }
  memcpy(psi, r, sizeof(rc) * S);
  free(r);
  return 0;
}

/* toffoli-gate (gate.clj:837-900). */
static int ref_toffoli(rc* psi, int n, int c1, int c2, int t)
{
  const uint64_t S = 1ull << n;
  rc* r = (rc*)calloc(S, sizeof(rc));
  if (!r) return -2;
  const int b1 = n - 1 - c1, b2 = n - 1 - c2, bt = n - 1 - t;
  for (uint64_t i = 0; i < S; ++i) {
    uint64_t dst = (BIT(i, b1) && BIT(i, b2)) ? (i ^ (1ull << bt)) : i;
    r[dst] = cadd(r[dst], psi[i]);
  }
  memcpy(psi, r, sizeof(rc) * S);
  free(r);
  return 0;
}

/* fredkin-gate (gate.clj:902-969). */
static int ref_fredkin(rc* psi, int n, int c, int t1, int t2)
{
  const uint64_t S = 1ull << n;
  rc* r = (rc*)calloc(S, sizeof(rc));
  if (!r) return -2;
  const int bc = n - 1 - c, b1 = n - 1 - t1, b2 = n - 1 - t2;
  for (uint64_t i = 0; i < S; ++i) {
    uint64_t dst = i;
    if (BIT(i, bc)) {
      uint64_t x1 = BIT(i, b1), x2 = BIT(i, b2);
      if (x1 != x2) dst = (i & ~(1ull << b1) & ~(1ull << b2)) | (x2 << b1) | (x1 << b2);
    }
    r[dst] = cadd(r[dst], psi[i]);
  }
  memcpy(psi, r, sizeof(rc) * S);
  free(r);
  return 0;
}

/* rydberg-cphase-gate (gate.clj:1003-1048): e^{i phi} where control and target are 1. */
static void ref_cphase(rc* psi, int n, int c, int t, double phi)
{
  const uint64_t S = 1ull << n;
  const int bc = n - 1 - c, bt = n - 1 - t;
  const rc ph = mk(cos(phi), sin(phi));
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)S; ++ii) {
    uint64_t i = (uint64_t)ii;
    if (BIT(i, bc) && BIT(i, bt)) psi[i] = cmul(psi[i], ph);
  }
}

/* rydberg-blockade-gate (gate.clj:1050-1095): e^{i phi} where exactly one of
 * the selected qubits is 1. */
static void ref_blockade(rc* psi, int n, uint64_t qmask, double phi)
{
  const uint64_t S = 1ull << n;
  uint64_t bmask = 0;
  for (int q = 0; q < n; ++q) if (BIT(qmask, q)) bmask |= 1ull << (n - 1 - q);
  const rc ph = mk(cos(phi), sin(phi));
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)S; ++ii) {
    if (__builtin_popcountll((uint64_t)ii & bmask) == 1) psi[ii] = cmul(psi[ii], ph);
  }
}

/* Counter-based uniform shared with libqsv (splitmix64). */
double ref_uniform(uint64_t seed, uint64_t k)
{
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* measure-specific-qubits (state.clj:915-1014) with a given uniform u.
 * Outcome bit j = j-th measured qubit in ascending qubit order. */
int ref_measure_qubits(double* amps, int n, uint64_t qmask, double u, uint64_t* outcome, double* probs_out)
{
  rc* psi = (rc*)amps;
  const uint64_t S = 1ull << n;
  int qs[64], m = 0;
  for (int q = 0; q < n; ++q) if (BIT(qmask, q)) qs[m++] = q;
  if (m == 0 || m > 24) return -1;
  const uint64_t no = 1ull << m;
  double* pr = (double*)calloc(no, sizeof(double));
  if (!pr) return -2;
  for (uint64_t i = 0; i < S; ++i) {
    uint64_t o = 0;
    for (int j = 0; j < m; ++j) o |= BIT(i, n - 1 - qs[j]) << j;
    pr[o] += psi[i].re * psi[i].re + psi[i].im * psi[i].im;
  }
  double total = 0;
  for (uint64_t o = 0; o < no; ++o) total += pr[o];
  const double r = u * total;
  double cum = 0;
  uint64_t sel = 0;
  for (uint64_t o = 0; o < no; ++o) { cum += pr[o]; if (cum < r) sel = o + 1; else break; }
  if (sel > no - 1) sel = no - 1;
  const double p = pr[sel];
  const double f = p > 0 ? 1.0 / sqrt(p) : 1.0;
  for (uint64_t i = 0; i < S; ++i) {
    uint64_t o = 0;
    for (int j = 0; j < m; ++j) o |= BIT(i, n - 1 - qs[j]) << j;
    if (o == sel) psi[i] = cmul(psi[i], mk(f, 0));
    else psi[i] = mk(0, 0);
  }
  if (outcome) *outcome = sel;
  if (probs_out) memcpy(probs_out, pr, sizeof(double) * no);
  free(pr);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* The dispatcher: apply-gate-to-state (circuit.clj:933-1072) over one op.    */
/* ------------------------------------------------------------------------ */
static int apply_1q(rc* psi, int n, const rc g[4], int q, int literal)
{
  if (q < 0 || q >= n) return -1;
  if (literal) return literal_1q(psi, n, g, q);
  pair_1q(psi, n, g, q);
  return 0;
}

static int apply_ctl(rc* psi, int n, int c, int t, const rc g[4], int literal)
{
  if (c < 0 || c >= n || t < 0 || t >= n || c == t) return -1;
  if (literal) return literal_controlled(psi, n, c, t, g);
  pair_controlled(psi, n, c, t, g);
  return 0;
}

int ref_apply_op(double* amps, int n, const qsv_op* op, uint32_t flags, int literal, uint64_t seed, uint64_t* measure_k)
{
  rc* psi = (rc*)amps;
  double md[8];
  rc g[4];
  const int k = op->kind;
  switch (k) {
  case QSV_OP_X: case QSV_OP_Y: case QSV_OP_Z: case QSV_OP_H: case QSV_OP_S: case QSV_OP_SDG:
  case QSV_OP_T: case QSV_OP_TDG: case QSV_OP_RX: case QSV_OP_RY: case QSV_OP_RZ: case QSV_OP_PHASE:
    ref_gate_matrix(k, op->angle, md); load_m(md, g);
    return apply_1q(psi, n, g, op->q[0], literal);
  case QSV_OP_U:
    load_m(op->m, g);
    return apply_1q(psi, n, g, op->q[0], literal);
  case QSV_OP_CNOT: case QSV_OP_CZ: case QSV_OP_CY: case QSV_OP_RYDBERG_CZ:
  case QSV_OP_CRX: case QSV_OP_CRY: case QSV_OP_CRZ: {
    int gk = (k == QSV_OP_CNOT) ? QSV_OP_X : (k == QSV_OP_CZ || k == QSV_OP_RYDBERG_CZ) ? QSV_OP_Z
           : (k == QSV_OP_CY) ? QSV_OP_Y : (k == QSV_OP_CRX) ? QSV_OP_RX : (k == QSV_OP_CRY) ? QSV_OP_RY : QSV_OP_RZ;
    ref_gate_matrix(gk, op->angle, md); load_m(md, g);
    return apply_ctl(psi, n, op->q[0], op->q[1], g, literal);
  }
  case QSV_OP_CU: {
    /* standard controlled-U: pass the transpose so that the reference-style
     * update (which applies g^T) yields U itself. */
    rc u[4]; load_m(op->m, u);
    g[0] = u[0]; g[1] = u[2]; g[2] = u[1]; g[3] = u[3];
    return apply_ctl(psi, n, op->q[0], op->q[1], g, literal);
  }
  case QSV_OP_SWAP: case QSV_OP_ISWAP: {
    int q1 = op->q[0], q2 = op->q[1];
    if (q1 < 0 || q2 < 0 || q1 >= n || q2 >= n || q1 == q2) return -1;
    return ref_swap_like(psi, n, q1, q2, k == QSV_OP_ISWAP, (flags & QSV_FLAG_MSB_SWAP) != 0);
  }
  case QSV_OP_TOFFOLI: {
    int a = op->q[0], b = op->q[1], c = op->q[2];
    if (a < 0 || b < 0 || c < 0 || a >= n || b >= n || c >= n || a == b || a == c || b == c) return -1;
    return ref_toffoli(psi, n, a, b, c);
  }
  case QSV_OP_FREDKIN: {
    int a = op->q[0], b = op->q[1], c = op->q[2];
    if (a < 0 || b < 0 || c < 0 || a >= n || b >= n || c >= n || a == b || a == c || b == c) return -1;
    return ref_fredkin(psi, n, a, b, c);
  }
  case QSV_OP_RYDBERG_CPHASE: {
    int c = op->q[0], t = op->q[1];
    if (c < 0 || t < 0 || c >= n || t >= n) return -1;
    ref_cphase(psi, n, c, t, op->angle);
    return 0;
  }
  case QSV_OP_RYDBERG_BLOCKADE:
    if (op->mask == 0 || (n < 64 && (op->mask >> n) != 0)) return -1;
    ref_blockade(psi, n, op->mask, op->angle);
    return 0;
  case QSV_OP_GLOBAL_RX: case QSV_OP_GLOBAL_RY: case QSV_OP_GLOBAL_RZ: case QSV_OP_GLOBAL_H:
  case QSV_OP_GLOBAL_X: case QSV_OP_GLOBAL_Y: case QSV_OP_GLOBAL_Z: {
    /* global-*-gate (gate.clj:1098-1253): the 1-qubit gate on qubits 0..n-1 in
     * turn; global-x/y/z are RX/RY/RZ(pi). */
    int gk; double ang = op->angle;
    switch (k) {
    case QSV_OP_GLOBAL_RX: gk = QSV_OP_RX; break;
    case QSV_OP_GLOBAL_RY: gk = QSV_OP_RY; break;
    case QSV_OP_GLOBAL_RZ: gk = QSV_OP_RZ; break;
    case QSV_OP_GLOBAL_H: gk = QSV_OP_H; break;
    case QSV_OP_GLOBAL_X: gk = QSV_OP_RX; ang = M_PI; break;
    case QSV_OP_GLOBAL_Y: gk = QSV_OP_RY; ang = M_PI; break;
    default: gk = QSV_OP_RZ; ang = M_PI; break;
    }
    ref_gate_matrix(gk, ang, md); load_m(md, g);
    for (int q = 0; q < n; ++q) { int rcode = apply_1q(psi, n, g, q, literal); if (rcode) return rcode; }
    return 0;
  }
  case QSV_OP_MEASURE: {
    if (op->mask == 0 || (n < 64 && (op->mask >> n) != 0)) return -1;
    double u = ref_uniform(seed, (*measure_k)++);
    return ref_measure_qubits(amps, n, op->mask, u, NULL, NULL);
  }
  default:
    return -3;  /* "Unknown gate type" */
  }
}

int ref_apply_program(double* amps, int n, const qsv_op* ops, uint64_t n_ops, uint32_t flags, int literal, uint64_t seed)
{
  uint64_t mk_ = 0;
  for (uint64_t i = 0; i < n_ops; ++i) {
    int r = ref_apply_op(amps, n, &ops[i], flags, literal, seed, &mk_);
    if (r) return r;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Result extraction.                                                         */
/* ------------------------------------------------------------------------ */
void ref_probabilities(const double* amps, int n, double* out)   /* state.clj:1071-1092 */
{
  const uint64_t S = 1ull << n;
  for (uint64_t i = 0; i < S; ++i) out[i] = amps[2 * i] * amps[2 * i] + amps[2 * i + 1] * amps[2 * i + 1];
}

double ref_norm2(const double* amps, int n)
{
  const uint64_t S = 1ull << n;
  double t = 0;
  for (uint64_t i = 0; i < S; ++i) t += amps[2 * i] * amps[2 * i] + amps[2 * i + 1] * amps[2 * i + 1];
  return t;
}

/* measure-state (state.clj:869-913): outcome = number of running sums < u*total. */
uint64_t ref_measure_state(const double* amps, int n, double u)
{
  const uint64_t S = 1ull << n;
  double total = ref_norm2(amps, n);
  double r = u * total, cum = 0;
  uint64_t cnt = 0;
  for (uint64_t i = 0; i < S; ++i) {
    cum += amps[2 * i] * amps[2 * i] + amps[2 * i + 1] * amps[2 * i + 1];
    if (cum < r) cnt = i + 1; else break;
  }
  return cnt < S ? cnt : S - 1;
}

void ref_sample(const double* amps, int n, uint64_t shots, uint64_t seed, uint64_t* out)
{
  for (uint64_t k = 0; k < shots; ++k) out[k] = ref_measure_state(amps, n, ref_uniform(seed, k));
}

/* The same shots as ref_sample for large states: the running sums of
 * measure-state (state.clj:869-913) are formed once, in the same order, into
 * `prefix` (2^n doubles, caller-provided), and each shot finds the first
 * running sum >= u * total by binary search, which is the count of running
 * sums below it because the sums never decrease. */
void ref_sample_prefix(const double* amps, int n, uint64_t shots, uint64_t seed, double* prefix, uint64_t* out)
{
  const uint64_t S = 1ull << n;
  double cum = 0;
  for (uint64_t i = 0; i < S; ++i) {
    cum += amps[2 * i] * amps[2 * i] + amps[2 * i + 1] * amps[2 * i + 1];
    prefix[i] = cum;
  }
  const double total = ref_norm2(amps, n);
  for (uint64_t k = 0; k < shots; ++k) {
    const double r = ref_uniform(seed, k) * total;
    uint64_t lo = 0, hi = S;   /* first index with prefix >= r, S if none */
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (prefix[mid] < r) lo = mid + 1; else hi = mid;
    }
    out[k] = lo < S ? lo : S - 1;
  }
}

/* pauli-string expectation by the Kronecker definition (observables.clj:216-251):
 * (P psi)_i = sum_j P_ij psi_j with P = sigma_{s[0]} ⊗ ... ⊗ sigma_{s[n-1]},
 * evaluated per row (each row has a single non-zero). */
double ref_pauli_expectation(const double* amps, int n, const char* s)
{
  const rc* psi = (const rc*)amps;
  const uint64_t S = 1ull << n;
  rc acc = { 0, 0 };
  for (uint64_t i = 0; i < S; ++i) {
    uint64_t j = i;
    rc coef = { 1, 0 };
    for (int q = 0; q < n; ++q) {
      int b = n - 1 - q;
      uint64_t bi = BIT(i, b);
      switch (s[q]) {
      case 'I': break;
      case 'X': j ^= 1ull << b; break;
      case 'Y': j ^= 1ull << b; coef = cmul(coef, bi ? mk(0, 1) : mk(0, -1)); break;  /* Y = [[0,-i],[i,0]] */
      case 'Z': if (bi) coef = mk(-coef.re, -coef.im); break;
      default: return NAN;
      }
    }
    rc ppsi = cmul(coef, psi[j]);
    rc conj_i = { psi[i].re, -psi[i].im };
    acc = cadd(acc, cmul(conj_i, ppsi));
  }
  return acc.re;
}

/* Dense form of the same: build the 2^n x 2^n Kronecker product of Paulis and
 * do mat-vec + inner product, exactly as the reference (n <= 10). */
double ref_pauli_expectation_literal(const double* amps, int n, const char* s)
{
  if (n > 10) return NAN;
  const rc* psi = (const rc*)amps;
  const uint64_t S = 1ull << n;
  rc* K = (rc*)malloc(sizeof(rc) * S * S);
  rc* T = (rc*)malloc(sizeof(rc) * S * S);
  rc P[4][4] = { { {1,0},{0,0},{0,0},{1,0} }, { {0,0},{1,0},{1,0},{0,0} },
                 { {0,0},{0,-1},{0,1},{0,0} }, { {1,0},{0,0},{0,0},{-1,0} } };
  #define PIDX(c) ((c) == 'I' ? 0 : (c) == 'X' ? 1 : (c) == 'Y' ? 2 : 3)
  uint64_t dim = 2;
  for (int k = 0; k < 4; ++k) K[k] = P[PIDX(s[0])][k];
  for (int f = 1; f < n; ++f) {
    const rc* F = P[PIDX(s[f])];
    uint64_t nd = dim * 2;
    for (uint64_t i = 0; i < dim; ++i)
      for (uint64_t j = 0; j < dim; ++j)
        for (int k = 0; k < 2; ++k)
          for (int l = 0; l < 2; ++l)
            T[(i * 2 + k) * nd + (j * 2 + l)] = cmul(K[i * dim + j], F[k * 2 + l]);
    rc* t = K; K = T; T = t;
    dim = nd;
  }
  rc acc = { 0, 0 };
  for (uint64_t i = 0; i < S; ++i) {
    rc row = { 0, 0 };
    for (uint64_t j = 0; j < S; ++j) row = cadd(row, cmul(K[i * S + j], psi[j]));
    rc ci = { psi[i].re, -psi[i].im };
    acc = cadd(acc, cmul(ci, row));
  }
  free(K); free(T);
  #undef PIDX
  return acc.re;
}

/* <a|b> (state.clj:1164-1185 returns its modulus). */
void ref_inner(const double* a, const double* b, int n, double out[2])
{
  const uint64_t S = 1ull << n;
  double re = 0, im = 0;
  for (uint64_t i = 0; i < S; ++i) {
    double ar = a[2 * i], ai = -a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
    re += ar * br - ai * bi;
    im += ar * bi + ai * br;
  }
  out[0] = re; out[1] = im;
}

/* <psi| I..O_q..I |psi> for a 1-qubit observable O (result.clj:254-288). */
void ref_expect_1q(const double* amps, int n, int q, const double m[8], double out[2])
{
  const rc* psi = (const rc*)amps;
  rc o[4]; load_m(m, o);
  const int b = n - 1 - q;
  const uint64_t half = 1ull << (n - 1), tb = 1ull << b, lo = tb - 1;
  double re = 0, im = 0;
  for (uint64_t p = 0; p < half; ++p) {
    uint64_t i0 = ((p & ~lo) << 1) | (p & lo), i1 = i0 | tb;
    rc a0 = psi[i0], a1 = psi[i1];
    rc y0 = cadd(cmul(o[0], a0), cmul(o[1], a1)), y1 = cadd(cmul(o[2], a0), cmul(o[3], a1));
    rc c0 = { a0.re, -a0.im }, c1 = { a1.re, -a1.im };
    rc t = cadd(cmul(c0, y0), cmul(c1, y1));
    re += t.re; im += t.im;
  }
  out[0] = re; out[1] = im;
}

/* Exact state of n qubits: |0..0> into amps (zero-state, state.clj:251-284). */
void ref_zero_state(double* amps, int n)
{
  const uint64_t S = 1ull << n;
  memset(amps, 0, sizeof(double) * 2 * S);
  amps[0] = 1.0;
}

int ref_literal_max_qubits(void) { return LITERAL_MAX_QUBITS; }
