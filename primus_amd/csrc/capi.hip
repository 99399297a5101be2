// capi.hip -- the extern "C" boundary declared in include/primus_hip.h.
//
// Table construction happens once on the host (exact 128-bit arithmetic) and
// uploads the Shoup twiddle tables, which then stay resident in HBM for the
// table's lifetime.  Transforms run on the GPU only: there is no host
// fallback, and a missing GPU is reported as PH_ERR_DEVICE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/primus_hip.h"
#include "arith.hpp"
#include "capi_internal.hpp"
#include "fhe_launch.hpp"
#include "ntt_launch.hpp"
#include "rns_launch.hpp"

using namespace ph;

struct ph_ntt_table {
    uint32_t log_n;
    uint64_t n, q, root, inv_root, inv_n, inv_n_q;
    std::vector<Shoup> fwd, inv, ord;  // host copies (accessors)
    Shoup *d_fwd = nullptr, *d_inv = nullptr, *d_ord = nullptr;
    ShoupF *d_fwd_f = nullptr, *d_inv_f = nullptr;  // binary64 twiddles (q < 2^50)
    ShoupF *d_perm[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [loge - 3][fwd, inv], register order
    bool f64 = false;                                // canonical ops may use F64Arith
    bool wide = false;                               // canonical ops may use IntCanonArith (q < 2^60)
    DevTable dev{};                    // device view (pointers into d_*)
    DevTable *d_tab = nullptr;         // the view itself in device memory (L = 1)
};

struct ph_dcrt_table {
    uint32_t log_n, L;
    bool f64 = false;
    bool wide = false;  // every modulus below 2^60
    bool owns = true;  // false for a view over tables owned elsewhere
    std::vector<ph_ntt_table *> tables;
    DevTable *d_tabs = nullptr;
};

struct ph_operand {
    uint32_t L;
    uint64_t n;
    Shoup *d = nullptr;      // ShoupFactor form (integer path)
    ShoupF *d_f[2] = {nullptr, nullptr};  // centred binary64 pairs {w, w/q} (binary64 path), register
                                          // order of 8 / 16 coefficients per thread
};

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(PH_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define PH_HIP(call)                                      \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)


void free_table(ph_ntt_table *t) {
    if (!t) return;
    if (t->d_fwd) (void)hipFree(t->d_fwd);
    if (t->d_inv) (void)hipFree(t->d_inv);
    if (t->d_ord) (void)hipFree(t->d_ord);
    if (t->d_fwd_f) (void)hipFree(t->d_fwd_f);
    if (t->d_inv_f) (void)hipFree(t->d_inv_f);
    for (auto &row : t->d_perm)
        for (ShoupF *p : row)
            if (p) (void)hipFree(p);
    if (t->d_tab) (void)hipFree(t->d_tab);
    delete t;
}

int upload(void **dst, const void *src, size_t bytes) {
    PH_HIP(hipMalloc(dst, bytes));
    PH_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return PH_OK;
}


// UintNttTable::new, primus_ntt/src/ntt/primitive.rs:114-185.
int build_table(uint32_t log_n, uint64_t q, uint32_t flags, ph_ntt_table **out) {
    if (q <= 1 || (64 - __builtin_clzll(q)) >= 63)
        return fail(PH_ERR_INVALID, "modulus must satisfy 1 < q < 2^62 (BarrettModulus::new)");
    if (log_n < kMinLogN || log_n > kMaxLogN)
        return fail(PH_ERR_UNSUPPORTED, "log_n must be in [1, 17]");
    HostTable h;
    const int st = host_table(log_n, q, h);
    if (st == PH_ERR_NO_PRIMITIVE_ROOT) return fail(st, "no primitive 2N-th root of unity modulo q");
    if (st == PH_ERR_DEGREE_TOO_LARGE) return fail(st, "degree must be less than the modulus");
    if (st != PH_OK) return fail(st, "host allocation failed");
    const uint64_t n = 1ull << log_n;
    ph_ntt_table *t = new (std::nothrow) ph_ntt_table();
    if (!t) return fail(PH_ERR_NO_MEMORY, "host allocation failed");
    t->log_n = log_n;
    t->n = n;
    t->q = q;
    t->root = h.root;
    t->inv_root = h.inv_root;
    t->ord = std::move(h.ord);
    t->fwd = std::move(h.fwd);
    t->inv = std::move(h.inv);
    t->inv_n = h.inv_n;
    t->inv_n_q = h.inv_n_q;

    int rc = upload((void **)&t->d_fwd, t->fwd.data(), n * sizeof(Shoup));
    if (rc == PH_OK) rc = upload((void **)&t->d_inv, t->inv.data(), n * sizeof(Shoup));
    if (rc == PH_OK) rc = upload((void **)&t->d_ord, t->ord.data(), 2 * n * sizeof(Shoup));
    if (rc != PH_OK) {
        free_table(t);
        return rc;
    }
    DevTable &d = t->dev;
    d.q = q;
    d.q2 = q << 1;
    d.q4 = q << 2;  // meaningful for q < 2^62
    d.q8 = q << 3;  // meaningful for q < 2^61; IntCanonArith uses both for q < 2^60
    d.inv_n = t->inv_n;
    d.inv_n_q = t->inv_n_q;
    const Shoup last = t->inv[n - 1];
    d.sr = h.sr;
    d.sr_q = h.sr_q;
    d.br0 = h.br0;
    d.br1 = h.br1;
    d.bar_mu = h.bar_mu;
    d.bar_sh = h.bar_sh;
    d.fwd = t->d_fwd;
    d.inv = t->d_inv;
    d.ord = t->d_ord;
    d.log_n = log_n;
    t->f64 = q < kF64MaxModulus && !(flags & PH_TABLE_INTEGER_ONLY);
    t->wide = q < kCanonIntMaxModulus;
    if (t->f64) {
        std::vector<ShoupF> ff(n), fi(n);
        for (uint64_t k = 0; k < n; ++k) {
            ff[k] = mkf(t->fwd[k], q);
            fi[k] = mkf(t->inv[k], q);
        }
        rc = upload((void **)&t->d_fwd_f, ff.data(), n * sizeof(ShoupF));
        if (rc == PH_OK) rc = upload((void **)&t->d_inv_f, fi.data(), n * sizeof(ShoupF));
        // register-order copies for the single-pass launch shapes (8 and 16
        // coefficients per thread)
        for (int loge = 3; loge <= 4 && rc == PH_OK && log_n <= kMaxSinglePassLogN; ++loge) {
            std::vector<ShoupF> pf, pi;
            register_order(ff, fi, log_n, loge, pf, pi);
            rc = upload((void **)&t->d_perm[loge - 3][0], pf.data(), n * sizeof(ShoupF));
            if (rc == PH_OK) rc = upload((void **)&t->d_perm[loge - 3][1], pi.data(), n * sizeof(ShoupF));
        }
        if (rc != PH_OK) {
            free_table(t);
            return rc;
        }
        d.qd = (double)q;
        d.qinv = 1.0 / (double)q;
        d.q_2p52 = (double)q + 0x1p52;
        ShoupF in = mkf(Shoup{t->inv_n, t->inv_n_q}, q);
        ShoupF sr = mkf(Shoup{d.sr, d.sr_q}, q);
        d.inv_n_f = in.w;
        d.inv_n_p = in.wp;
        d.sr_f = sr.w;
        d.sr_p = sr.wp;
        const ShoupF lastf = mkf(last, q);
        d.last_f = lastf.w;
        d.last_p = lastf.wp;
        d.fwd_f = t->d_fwd_f;
        d.inv_f = t->d_inv_f;
        for (int e = 0; e < 2; ++e) {
            d.fwd_p[e] = t->d_perm[e][0];
            d.inv_p[e] = t->d_perm[e][1];
        }
    }
    rc = upload((void **)&t->d_tab, &d, sizeof(DevTable));
    if (rc != PH_OK) {
        free_table(t);
        return rc;
    }
    *out = t;
    return PH_OK;
}

// One pinned staging buffer and one stream per host thread for the host-slice
// entry points.
struct Staging {
    void *buf = nullptr;
    size_t cap = 0;
    hipStream_t stream = nullptr;
};
thread_local Staging g_staging;

int staging(size_t bytes, void **buf, hipStream_t *stream) {
    Staging &s = g_staging;
    if (!s.stream) PH_HIP(hipStreamCreate(&s.stream));
    if (s.cap < bytes) {
        if (s.buf) PH_HIP(hipFree(s.buf));
        s.buf = nullptr;
        s.cap = 0;
        PH_HIP(hipMalloc(&s.buf, bytes));
        s.cap = bytes;
    }
    *buf = s.buf;
    *stream = s.stream;
    return PH_OK;
}

// Autotuned launch shape per (log_n, op, arithmetic): log2 of the coefficients
// each thread holds in single-pass kernels (0 = the default).  Written by
// ph_tuning_set / ph_autotune / ph_tuning_load, read by every launch.
std::atomic<int> g_tuning[kMaxLogN + 1][kOpCount][2];

// The arithmetic a table's launches may use (kArithF64 / kArithCanonInt bits).
template <class Tab>
int arith_of(const Tab *t) {
    return (t->f64 ? kArithF64 : 0) | (t->wide ? kArithCanonInt : 0);
}

int tuning_of(uint32_t log_n, Op op, bool f64) {
    if (log_n > kMaxLogN) return 0;
    return g_tuning[log_n][(int)op][f64 ? 1 : 0].load(std::memory_order_relaxed);
}

// limb_of: per limb-polynomial table index (poly-multiply), or null for i % L.
// caller_scratch: device memory of scratch_bytes(op, log_n, count) or more to
// use instead of an allocation, or null.
int run_op(Op op, const DevTable *d_tabs, uint32_t L, uint32_t log_n, int arith, uint64_t *d_data,
           const uint64_t *d_data2, const ph_operand *operand, uint64_t count, hipStream_t s,
           const uint8_t *limb_of = nullptr, void *caller_scratch = nullptr) {
    if (count == 0) return PH_OK;
    KArgs a{};
    a.limb_of = limb_of;
    a.data = d_data;
    a.data2 = d_data2;
    a.operand = operand ? operand->d : nullptr;
    a.operand_f[0] = operand ? operand->d_f[0] : nullptr;
    a.operand_f[1] = operand ? operand->d_f[1] : nullptr;
    if (op == Op::FwdMulInv && (!a.operand_f[0] || !a.operand_f[1])) arith &= ~kArithF64;
    a.tabs = d_tabs;
    a.L = L;
    void *scratch = caller_scratch;
    const bool own = needs_scratch(op, log_n) && !scratch;
    if (own) PH_HIP(hipMallocAsync(&scratch, (count << log_n) * 8, s));
    const bool f64 = (arith & kArithF64) != 0 && op != Op::FwdLazy;
    hipError_t e = launch_transform(op, a, log_n, count, scratch, arith, tuning_of(log_n, op, f64), s);
    if (own) (void)hipFreeAsync(scratch, s);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return PH_OK;
}

// Host slice -> device -> op -> host slice, synchronous.
int host_roundtrip(Op op, const DevTable *d_tabs, uint32_t L, uint32_t log_n, int arith, uint64_t *poly,
                   size_t len) {
    void *buf;
    hipStream_t s;
    int rc = staging(len * 8, &buf, &s);
    if (rc != PH_OK) return rc;
    PH_HIP(hipMemcpyAsync(buf, poly, len * 8, hipMemcpyHostToDevice, s));
    rc = run_op(op, d_tabs, L, log_n, arith, (uint64_t *)buf, nullptr, nullptr, len >> log_n, s);
    if (rc != PH_OK) return rc;
    PH_HIP(hipMemcpyAsync(poly, buf, len * 8, hipMemcpyDeviceToHost, s));
    PH_HIP(hipStreamSynchronize(s));
    return PH_OK;
}

int check_slice(const void *table, const void *p, size_t len, size_t want) {
    if (!table || (!p && len)) return fail(PH_ERR_INVALID, "null table or slice");
    if (len != want) return fail(PH_ERR_INVALID, "slice length does not match the table");
    return PH_OK;
}

int monomial(const ph_ntt_table *t, uint64_t coeff, size_t degree, uint64_t *values, size_t len,
             int kind) {
    int rc = check_slice(t, values, len, t ? t->n : 0);
    if (rc != PH_OK) return rc;
    void *buf;
    hipStream_t s;
    rc = staging(len * 8, &buf, &s);
    if (rc != PH_OK) return rc;
    hipError_t e = launch_monomial(t->d_tab, t->log_n, (uint64_t *)buf, coeff, (uint64_t)degree, kind, s);
    if (e != hipSuccess) return hip_fail(e, "monomial kernel");
    PH_HIP(hipMemcpyAsync(values, buf, len * 8, hipMemcpyDeviceToHost, s));
    PH_HIP(hipStreamSynchronize(s));
    return PH_OK;
}

}  // namespace

extern "C" {

const char *ph_status_string(int status) {
    switch (status) {
        case PH_OK: return "ok";
        case PH_ERR_NO_PRIMITIVE_ROOT: return "no primitive root";
        case PH_ERR_DEGREE_CONVERSION: return "degree conversion";
        case PH_ERR_DEGREE_TOO_LARGE: return "degree too large";
        case PH_ERR_NTT_TABLE: return "ntt table";
        case PH_ERR_INVALID: return "invalid argument";
        case PH_ERR_DEVICE: return "device error";
        case PH_ERR_UNSUPPORTED: return "unsupported ring degree";
        case PH_ERR_NO_MEMORY: return "out of memory";
        case PH_ERR_CANCELLED: return "cancelled";
    }
    return "unknown status";
}

const char *ph_last_error(void) { return g_last_error.c_str(); }

int ph_debug_clock_report(uint64_t *entries, uint32_t *counts, int clear) {
    const hipError_t e = clock_probe_report(entries, counts, clear != 0);
    if (e == hipErrorNotSupported) return fail(PH_ERR_UNSUPPORTED, "library built without PH_CLOCK_PROBE");
    if (e != hipSuccess) return hip_fail(e, "clock-probe report");
    return PH_OK;
}

int ph_debug_range_report(uint64_t report[4], int clear) {
    if (!report) return fail(PH_ERR_INVALID, "null report");
#if PH_RANGE_CHECK
    uint64_t a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
    hipError_t e = range_report_ntt(a, clear != 0);
    if (e == hipSuccess) e = range_report_fhe(b, clear != 0);
    if (e != hipSuccess) return hip_fail(e, "range-check report");
    const uint64_t *first = a[0] ? a : b;
    report[0] = a[0] + b[0];
    for (int i = 1; i < 4; ++i) report[i] = first[i];
    return PH_OK;
#else
    (void)clear;
    for (int i = 0; i < 4; ++i) report[i] = 0;
    return fail(PH_ERR_UNSUPPORTED, "library built without PH_RANGE_CHECK");
#endif
}

int ph_ntt_table_new(uint32_t log_n, uint64_t modulus, ph_ntt_table **out) {
    return ph_ntt_table_new_ex(log_n, modulus, 0, out);
}

// The tuning table a deployment keeps in the file PH_TUNING_FILE names
// (ph_tuning_save's format), loaded once, when the first table is created.  A
// file that does not exist yet is no error (nothing has been tuned); one that
// holds an invalid entry fails every table creation, so that a damaged table
// is not silently ignored.
static int load_env_tuning() {
    static std::once_flag once;
    static int rc = PH_OK;
    std::call_once(once, [] {
        const char *path = getenv("PH_TUNING_FILE");
        if (!path || !*path) return;
        FILE *f = fopen(path, "r");
        if (!f) return;
        fclose(f);
        rc = ph_tuning_load(path);
    });
    return rc == PH_OK ? PH_OK : fail(rc, "the tuning file PH_TUNING_FILE names holds an invalid entry");
}

int ph_ntt_table_new_ex(uint32_t log_n, uint64_t modulus, uint32_t flags, ph_ntt_table **out) {
    if (!out) return fail(PH_ERR_INVALID, "null output");
    *out = nullptr;
    if (const int rc = load_env_tuning()) return rc;
    return build_table(log_n, modulus, flags, out);
}

int ph_ntt_table_arith(const ph_ntt_table *t) { return t ? (t->f64 ? 1 : 0) : -1; }

void ph_ntt_table_free(ph_ntt_table *table) { free_table(table); }

size_t ph_ntt_table_poly_length(const ph_ntt_table *t) { return t ? t->n : 0; }
uint32_t ph_ntt_table_log_n(const ph_ntt_table *t) { return t ? t->log_n : 0; }
uint64_t ph_ntt_table_modulus(const ph_ntt_table *t) { return t ? t->q : 0; }
uint64_t ph_ntt_table_root(const ph_ntt_table *t) { return t ? t->root : 0; }
uint64_t ph_ntt_table_inv_root(const ph_ntt_table *t) { return t ? t->inv_root : 0; }

int ph_ntt_table_inv_n(const ph_ntt_table *t, uint64_t *value, uint64_t *quotient) {
    if (!t) return fail(PH_ERR_INVALID, "null table");
    if (value) *value = t->inv_n;
    if (quotient) *quotient = t->inv_n_q;
    return PH_OK;
}

int ph_ntt_table_powers(const ph_ntt_table *t, int which, uint64_t *values, uint64_t *quotients) {
    if (!t || which < 0 || which > 2) return fail(PH_ERR_INVALID, "bad table or selector");
    const std::vector<Shoup> &v = which == 0 ? t->fwd : which == 1 ? t->inv : t->ord;
    for (size_t i = 0; i < v.size(); ++i) {
        if (values) values[i] = v[i].w;
        if (quotients) quotients[i] = v[i].wq;
    }
    return PH_OK;
}

int ph_lazy_transform_slice(const ph_ntt_table *t, uint64_t *poly, size_t len) {
    int rc = check_slice(t, poly, len, t ? t->n : 0);
    return rc != PH_OK ? rc : host_roundtrip(Op::FwdLazy, t->d_tab, 1, t->log_n, arith_of(t), poly, len);
}

int ph_transform_slice(const ph_ntt_table *t, uint64_t *poly, size_t len) {
    int rc = check_slice(t, poly, len, t ? t->n : 0);
    return rc != PH_OK ? rc : host_roundtrip(Op::FwdCanon, t->d_tab, 1, t->log_n, arith_of(t), poly, len);
}

int ph_lazy_inverse_transform_slice(const ph_ntt_table *t, uint64_t *values, size_t len) {
    int rc = check_slice(t, values, len, t ? t->n : 0);
    return rc != PH_OK ? rc : host_roundtrip(Op::Inv, t->d_tab, 1, t->log_n, arith_of(t), values, len);
}

int ph_inverse_transform_slice(const ph_ntt_table *t, uint64_t *values, size_t len) {
    int rc = check_slice(t, values, len, t ? t->n : 0);
    return rc != PH_OK ? rc : host_roundtrip(Op::Inv, t->d_tab, 1, t->log_n, arith_of(t), values, len);
}

int ph_hexl_compute_forward(const ph_ntt_table *t, uint64_t *operand, size_t len, uint64_t input_mod_factor,
                            uint64_t output_mod_factor) {
    if (input_mod_factor != 1 && input_mod_factor != 2 && input_mod_factor != 4)
        return fail(PH_ERR_INVALID, "input_mod_factor must be 1, 2 or 4");
    if (output_mod_factor != 1 && output_mod_factor != 4) return fail(PH_ERR_INVALID, "output_mod_factor must be 1 or 4");
    // both forward kernels take [0,4q)
    return output_mod_factor == 1 ? ph_transform_slice(t, operand, len) : ph_lazy_transform_slice(t, operand, len);
}

int ph_hexl_compute_inverse(const ph_ntt_table *t, uint64_t *operand, size_t len, uint64_t input_mod_factor,
                            uint64_t output_mod_factor) {
    if (input_mod_factor != 1 && input_mod_factor != 2) return fail(PH_ERR_INVALID, "input_mod_factor must be 1 or 2");
    if (output_mod_factor != 1 && output_mod_factor != 2) return fail(PH_ERR_INVALID, "output_mod_factor must be 1 or 2");
    // output_mod_factor 2: HEXL's radix-2 inverse, whose last stage leaves
    // lazy products in [0, 2q) (hexl/radix2.rs:358-374); 1: canonical
    if (output_mod_factor == 1) return ph_inverse_transform_slice(t, operand, len);
    const int rc = check_slice(t, operand, len, t ? t->n : 0);
    return rc != PH_OK ? rc : host_roundtrip(Op::InvHexlLazy, t->d_tab, 1, t->log_n, arith_of(t), operand, len);
}

int ph_transform_monomial(const ph_ntt_table *t, uint64_t coeff, size_t degree, uint64_t *values,
                          size_t len) {
    return monomial(t, coeff, degree, values, len, 0);
}

int ph_transform_coeff_one_monomial(const ph_ntt_table *t, size_t degree, uint64_t *values, size_t len) {
    return monomial(t, 0, degree, values, len, 1);
}

int ph_transform_coeff_minus_one_monomial(const ph_ntt_table *t, size_t degree, uint64_t *values,
                                          size_t len) {
    return monomial(t, 0, degree, values, len, 2);
}

int ph_dcrt_table_new(uint32_t log_n, const uint64_t *moduli, uint32_t moduli_count, ph_dcrt_table **out) {
    return ph_dcrt_table_new_ex(log_n, moduli, moduli_count, 0, out);
}

int ph_dcrt_table_new_ex(uint32_t log_n, const uint64_t *moduli, uint32_t moduli_count, uint32_t flags,
                         ph_dcrt_table **out) {
    if (!out || !moduli || moduli_count == 0) return fail(PH_ERR_INVALID, "null moduli or output");
    *out = nullptr;
    if (const int rc = load_env_tuning()) return rc;
    ph_dcrt_table *d = new (std::nothrow) ph_dcrt_table();
    if (!d) return fail(PH_ERR_NO_MEMORY, "host allocation failed");
    d->log_n = log_n;
    d->L = moduli_count;
    std::vector<DevTable> views;
    d->f64 = true;
    d->wide = true;
    for (uint32_t i = 0; i < moduli_count; ++i) {
        ph_ntt_table *t = nullptr;
        int rc = build_table(log_n, moduli[i], flags, &t);
        if (rc != PH_OK) {
            ph_dcrt_table_free(d);
            return rc;
        }
        d->tables.push_back(t);
        views.push_back(t->dev);
        d->f64 = d->f64 && t->f64;
        d->wide = d->wide && t->wide;
    }
    int rc = upload((void **)&d->d_tabs, views.data(), views.size() * sizeof(DevTable));
    if (rc != PH_OK) {
        ph_dcrt_table_free(d);
        return rc;
    }
    *out = d;
    return PH_OK;
}

int ph_dcrt_table_view(const ph_ntt_table *const *tables, uint32_t moduli_count, ph_dcrt_table **out) {
    if (!out || !tables || moduli_count == 0) return fail(PH_ERR_INVALID, "null tables or output");
    *out = nullptr;
    const uint32_t log_n = tables[0] ? tables[0]->log_n : 0;
    for (uint32_t i = 0; i < moduli_count; ++i)
        if (!tables[i] || tables[i]->log_n != log_n) return fail(PH_ERR_INVALID, "tables must share the ring degree");
    ph_dcrt_table *d = new (std::nothrow) ph_dcrt_table();
    if (!d) return fail(PH_ERR_NO_MEMORY, "host allocation failed");
    d->log_n = log_n;
    d->L = moduli_count;
    d->owns = false;
    d->f64 = true;
    d->wide = true;
    std::vector<DevTable> views;
    for (uint32_t i = 0; i < moduli_count; ++i) {
        d->tables.push_back(const_cast<ph_ntt_table *>(tables[i]));
        views.push_back(tables[i]->dev);
        d->f64 = d->f64 && tables[i]->f64;
        d->wide = d->wide && tables[i]->wide;
    }
    int rc = upload((void **)&d->d_tabs, views.data(), views.size() * sizeof(DevTable));
    if (rc != PH_OK) {
        ph_dcrt_table_free(d);
        return rc;
    }
    *out = d;
    return PH_OK;
}

void ph_dcrt_table_free(ph_dcrt_table *d) {
    if (!d) return;
    if (d->owns)
        for (ph_ntt_table *t : d->tables) free_table(t);
    if (d->d_tabs) (void)hipFree(d->d_tabs);
    delete d;
}

int ph_dcrt_table_arith(const ph_dcrt_table *d) { return d ? (d->f64 ? 1 : 0) : -1; }

size_t ph_dcrt_poly_length(const ph_dcrt_table *d) { return d ? (size_t)1 << d->log_n : 0; }
size_t ph_dcrt_moduli_count(const ph_dcrt_table *d) { return d ? d->L : 0; }
size_t ph_dcrt_crt_poly_length(const ph_dcrt_table *d) { return d ? ((size_t)d->L << d->log_n) : 0; }

const ph_ntt_table *ph_dcrt_ntt_table(const ph_dcrt_table *d, uint32_t i) {
    return (d && i < d->L) ? d->tables[i] : nullptr;
}

static int dcrt_slice(const ph_dcrt_table *d, uint64_t *poly, size_t len, Op op) {
    int rc = check_slice(d, poly, len, d ? ((size_t)d->L << d->log_n) : 0);
    return rc != PH_OK ? rc : host_roundtrip(op, d->d_tabs, d->L, d->log_n, arith_of(d), poly, len);
}

int ph_dcrt_lazy_transform_slice(const ph_dcrt_table *d, uint64_t *poly, size_t len) {
    return dcrt_slice(d, poly, len, Op::FwdLazy);
}
int ph_dcrt_transform_slice(const ph_dcrt_table *d, uint64_t *poly, size_t len) {
    return dcrt_slice(d, poly, len, Op::FwdCanon);
}
int ph_dcrt_lazy_inverse_transform_slice(const ph_dcrt_table *d, uint64_t *poly, size_t len) {
    return dcrt_slice(d, poly, len, Op::Inv);
}
int ph_dcrt_inverse_transform_slice(const ph_dcrt_table *d, uint64_t *poly, size_t len) {
    return dcrt_slice(d, poly, len, Op::Inv);
}

int ph_dcrt_transform_device(const ph_dcrt_table *d, uint64_t *d_polys, size_t count, int lazy, void *stream) {
    if (!d || (!d_polys && count)) return fail(PH_ERR_INVALID, "null table or data");
    return run_op(lazy ? Op::FwdLazy : Op::FwdCanon, d->d_tabs, d->L, d->log_n, arith_of(d), d_polys, nullptr, nullptr,
                  count, (hipStream_t)stream);
}

int ph_dcrt_inverse_transform_device(const ph_dcrt_table *d, uint64_t *d_polys, size_t count, void *stream) {
    if (!d || (!d_polys && count)) return fail(PH_ERR_INVALID, "null table or data");
    return run_op(Op::Inv, d->d_tabs, d->L, d->log_n, arith_of(d), d_polys, nullptr, nullptr, count, (hipStream_t)stream);
}

int ph_ntt_transform_device(const ph_ntt_table *t, uint64_t *d_polys, size_t count, int lazy, void *stream) {
    if (!t || (!d_polys && count)) return fail(PH_ERR_INVALID, "null table or data");
    return run_op(lazy ? Op::FwdLazy : Op::FwdCanon, t->d_tab, 1, t->log_n, arith_of(t), d_polys, nullptr, nullptr, count,
                  (hipStream_t)stream);
}

int ph_ntt_inverse_transform_device(const ph_ntt_table *t, uint64_t *d_polys, size_t count, void *stream) {
    if (!t || (!d_polys && count)) return fail(PH_ERR_INVALID, "null table or data");
    return run_op(Op::Inv, t->d_tab, 1, t->log_n, arith_of(t), d_polys, nullptr, nullptr, count, (hipStream_t)stream);
}

int ph_operand_new(const ph_dcrt_table *d, const uint64_t *host_values, ph_operand **out) {
    if (!d || !host_values || !out) return fail(PH_ERR_INVALID, "null argument");
    *out = nullptr;
    const uint64_t n = 1ull << d->log_n;
    std::vector<Shoup> h((size_t)d->L * n);

    for (uint32_t l = 0; l < d->L; ++l) {
        const uint64_t q = d->tables[l]->q;
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t v = host_values[l * n + i];
            if (v >= q) return fail(PH_ERR_INVALID, "operand values must be canonical (< q)");
            h[l * n + i] = mk(v, q);

        }
    }
    ph_operand *op = new (std::nothrow) ph_operand();
    if (!op) return fail(PH_ERR_NO_MEMORY, "host allocation failed");
    op->L = d->L;
    op->n = n;
    int rc = upload((void **)&op->d, h.data(), h.size() * sizeof(Shoup));
    if (rc == PH_OK && d->f64) {
        // the binary64 copies hold operand * N^-1: the fused kernel's inverse
        // then needs no N^-1 multiply on its last stage's sums
        std::vector<ShoupF> hf(h.size());
        for (uint32_t l = 0; l < d->L; ++l) {
            const uint64_t q = d->tables[l]->q, inv_n = d->tables[l]->inv_n;
            for (uint64_t i = 0; i < n; ++i) hf[l * n + i] = mkf(mk(h_mulmod(h[l * n + i].w, inv_n, q), q), q);
        }
        for (int e = 0; e < 2 && rc == PH_OK; ++e) {
            const std::vector<ShoupF> reg = operand_register_order(hf, d->L, d->log_n, e + 3);
            rc = upload((void **)&op->d_f[e], reg.data(), reg.size() * sizeof(ShoupF));
        }
    }
    if (rc != PH_OK) {
        ph_operand_free(op);
        return rc;
    }
    *out = op;
    return PH_OK;
}

void ph_operand_free(ph_operand *op) {
    if (!op) return;
    if (op->d) (void)hipFree(op->d);
    for (ShoupF *p : op->d_f)
        if (p) (void)hipFree(p);
    delete op;
}

int ph_dcrt_fwd_mul_inv_device(const ph_dcrt_table *d, uint64_t *d_polys, size_t count, const ph_operand *op,
                               void *stream) {
    if (!d || !op || (!d_polys && count)) return fail(PH_ERR_INVALID, "null argument");
    if (op->L != d->L || op->n != (1ull << d->log_n)) return fail(PH_ERR_INVALID, "operand shape mismatch");
    return run_op(Op::FwdMulInv, d->d_tabs, d->L, d->log_n, arith_of(d), d_polys, nullptr, op, count, (hipStream_t)stream);
}

int ph_dcrt_negacyclic_mul_device(const ph_dcrt_table *d, uint64_t *d_a, const uint64_t *d_b, size_t count,
                                  void *stream) {
    if (!d || ((!d_a || !d_b) && count)) return fail(PH_ERR_INVALID, "null argument");
    return run_op(Op::PolyMul, d->d_tabs, d->L, d->log_n, arith_of(d), d_a, d_b, nullptr, count, (hipStream_t)stream);
}

}  // extern "C"

namespace {

// Scratch of ph_dcrt_negacyclic_mul_device_noalloc: the column pass of the
// second operand for two-pass ring degrees.  Every GPU the library runs on
// holds its own copy for the life of the process.  In host code the name is
// only a stand-in for the symbol, never a device address: kernels and copies
// get the address hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocScratchWords = PH_NOALLOC_SCRATCH_BYTES / 8;
__device__ __align__(256) uint64_t scratch[kNoAllocScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up on the first call on
// that device and kept.
int noalloc_scratch(void **out) {
    int dev = -1;
    PH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(PH_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(PH_ERR_DEVICE, "no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PH_OK;
}

}  // namespace

extern "C" {

int ph_dcrt_negacyclic_mul_device_noalloc(const ph_dcrt_table *d, uint64_t *d_a, const uint64_t *d_b, size_t count,
                                          void *stream) {
    if (!d || ((!d_a || !d_b) && count)) return fail(PH_ERR_INVALID, "null argument");
    if (count == 0) return PH_OK;
    const uint64_t need = needs_scratch(Op::PolyMul, d->log_n) ? ((uint64_t)count << d->log_n) : 0;
    if (need > kNoAllocScratchWords)
        return fail(PH_ERR_INVALID, "count * N * 8 bytes exceeds the fixed scratch of PH_NOALLOC_SCRATCH_BYTES");
    void *scr = nullptr;
    if (need) {
        const int rc = noalloc_scratch(&scr);
        if (rc != PH_OK) return rc;
    }
    // every scratch word the column pass writes is read back by the chunk
    // pass of the same launch sequence, so nothing needs clearing
    return run_op(Op::PolyMul, d->d_tabs, d->L, d->log_n, arith_of(d), d_a, d_b, nullptr, count, (hipStream_t)stream,
                  nullptr, scr);
}

int ph_dcrt_negacyclic_mul_rows_device(const ph_dcrt_table *d, uint64_t *d_a, const uint64_t *d_b,
                                       const uint8_t *d_limb_of, size_t count, void *stream) {
    if (!d || ((!d_a || !d_b || !d_limb_of) && count)) return fail(PH_ERR_INVALID, "null argument");
    return run_op(Op::PolyMul, d->d_tabs, d->L, d->log_n, arith_of(d), d_a, d_b, nullptr, count, (hipStream_t)stream,
                  d_limb_of);
}

int ph_dcrt_mul_factor_device(const ph_dcrt_table *d, uint64_t *d_polys, size_t count, const uint64_t *factors,
                              uint32_t bit_shift, int lazy, void *stream) {
    if (!d || !factors || (!d_polys && count)) return fail(PH_ERR_INVALID, "null argument");
    if (bit_shift != 32 && bit_shift != 52 && bit_shift != 64)
        return fail(PH_ERR_INVALID, "bit_shift must be 32, 52 or 64 (MultiplyFactor::new)");
    if (d->L > kMulFactorMaxLimbs) return fail(PH_ERR_UNSUPPORTED, "at most 64 limbs");
    MulFactorArgs a{};
    a.data = d_polys;
    a.total = (uint64_t)count << d->log_n;
    a.L = d->L;
    a.log_n = d->log_n;
    a.lazy = lazy ? 1u : 0u;
    for (uint32_t i = 0; i < d->L; ++i) {
        const uint64_t q = d->tables[i]->q, c = factors[i];
        // MultiplyFactor::new asserts operand <= modulus; the shift bounds q
        // (multiply_mod_lazy's debug_assert, hexl/number_theory.rs:74-76)
        if (c > q) return fail(PH_ERR_INVALID, "factor must not exceed its limb modulus");
        if (bit_shift < 64 && q >= (1ull << bit_shift)) return fail(PH_ERR_INVALID, "modulus exceeds 2^bit_shift");
        a.factor[i] = c;
        a.barrett[i] = (uint64_t)((((u128)c) << bit_shift) / q);
        a.q[i] = q;
    }
    const hipError_t e = launch_mul_factor(a, bit_shift, (hipStream_t)stream);
    return e == hipSuccess ? PH_OK : hip_fail(e, "multiply-factor kernel");
}

int ph_dcrt_pointwise_mul_device(const ph_dcrt_table *d, const uint64_t *d_a, const uint64_t *d_b,
                                 uint64_t *d_out, size_t count, void *stream) {
    if (!d || ((!d_a || !d_b || !d_out) && count)) return fail(PH_ERR_INVALID, "null argument");
    hipError_t e = launch_pointwise(d->d_tabs, d->L, d->log_n, (uint64_t)count << d->log_n, d_a, d_b, d_out, 0,
                                    (hipStream_t)stream);
    return e == hipSuccess ? PH_OK : hip_fail(e, "pointwise kernel");
}

int ph_dcrt_pointwise_mul_add_device(const ph_dcrt_table *d, uint64_t *d_acc, const uint64_t *d_a,
                                     const uint64_t *d_b, size_t count, void *stream) {
    if (!d || ((!d_a || !d_b || !d_acc) && count)) return fail(PH_ERR_INVALID, "null argument");
    hipError_t e = launch_pointwise(d->d_tabs, d->L, d->log_n, (uint64_t)count << d->log_n, d_a, d_b, d_acc, 1,
                                    (hipStream_t)stream);
    return e == hipSuccess ? PH_OK : hip_fail(e, "pointwise kernel");
}

int ph_dcrt_automorphism_device(const ph_dcrt_table *d, const uint64_t *d_in, uint64_t *d_out, size_t count,
                                uint64_t degree, void *stream) {
    if (!d || ((!d_in || !d_out) && count)) return fail(PH_ERR_INVALID, "null argument");
    if ((degree & 1) == 0) return fail(PH_ERR_INVALID, "automorphism degree must be odd");
    if (count == 0) return PH_OK;
    const uint64_t bytes = (uint64_t)count << (d->log_n + 3);
    const uintptr_t a = (uintptr_t)d_in, b = (uintptr_t)d_out;
    if (a < b + bytes && b < a + bytes) return fail(PH_ERR_INVALID, "input and output overlap");
    const uint32_t k = (uint32_t)(degree & ((2ull << d->log_n) - 1));
    hipError_t e = launch_automorphism(d_in, d_out, d->log_n, k, count, (hipStream_t)stream);
    return e == hipSuccess ? PH_OK : hip_fail(e, "automorphism kernel");
}

// ---------------------------------------------------------------------------
// RNS base conversion
// ---------------------------------------------------------------------------

struct ph_base_converter {
    uint32_t in_count = 0, out_count = 0;
    void *d_mem = nullptr;  // every BaseConvDev array, one allocation
    BaseConvDev dev{};
};

void ph_base_converter_free(ph_base_converter *c) {
    if (!c) return;
    if (c->d_mem) (void)hipFree(c->d_mem);
    delete c;
}

static uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

int ph_base_converter_new(const uint64_t *ibase, uint32_t in_count, const uint64_t *obase, uint32_t out_count,
                          ph_base_converter **out) {
    if (!out || !ibase || !obase || in_count == 0 || out_count == 0) return fail(PH_ERR_INVALID, "bad arguments");
    *out = nullptr;
    if (in_count > kMaxRnsModuli || out_count > kMaxRnsModuli)
        return fail(PH_ERR_UNSUPPORTED, "at most 32 moduli per base");
    for (uint32_t i = 0; i < in_count; ++i) {
        if (ibase[i] <= 1 || ibase[i] >= (1ull << 62)) return fail(PH_ERR_INVALID, "modulus outside (1, 2^62)");
        for (uint32_t k = i + 1; k < in_count; ++k)
            if (gcd64(ibase[i], ibase[k]) != 1) return fail(PH_ERR_INVALID, "input moduli not pairwise coprime");
    }
    for (uint32_t j = 0; j < out_count; ++j)
        if (obase[j] <= 1 || obase[j] >= (1ull << 62)) return fail(PH_ERR_INVALID, "modulus outside (1, 2^62)");
    // host image of the device arrays, in BaseConvDev order
    const size_t I = in_count, J = out_count;
    std::vector<uint64_t> img(I * 5 + I + J * 3 + J * I + J);
    uint64_t *q = img.data(), *q_r0 = q + I, *q_r1 = q_r0 + I, *inv = q_r1 + I, *inv_q = inv + I;
    double *qd = (double *)(inv_q + I);
    uint64_t *p = inv_q + 2 * I, *p_r0 = p + J, *p_r1 = p_r0 + J, *m = p_r1 + J, *q_mod_p = m + J * I;
    for (size_t i = 0; i < I; ++i) {
        q[i] = ibase[i];
        h_barrett_ratio(q[i], &q_r0[i], &q_r1[i]);
        uint64_t punct = 1 % q[i];  // (Q / q_i) mod q_i, RNSBase::new (primus_rns/src/base.rs:66-82)
        for (size_t k = 0; k < I; ++k)
            if (k != i) punct = h_mulmod(punct, ibase[k] % q[i], q[i]);
        inv[i] = inv_mod(punct, q[i]);
        inv_q[i] = h_shoup_quotient(inv[i], q[i]);
        qd[i] = (double)q[i];
    }
    for (size_t j = 0; j < J; ++j) {
        p[j] = obase[j];
        h_barrett_ratio(p[j], &p_r0[j], &p_r1[j]);
        uint64_t all = 1 % p[j];
        for (size_t k = 0; k < I; ++k) all = h_mulmod(all, ibase[k] % p[j], p[j]);
        q_mod_p[j] = all;
        for (size_t i = 0; i < I; ++i) {  // BaseConverter::new (primus_rns/src/converter.rs:31-41)
            uint64_t punct = 1 % p[j];
            for (size_t k = 0; k < I; ++k)
                if (k != i) punct = h_mulmod(punct, ibase[k] % p[j], p[j]);
            m[j * I + i] = punct;
        }
    }
    // binary64 form when every modulus is below 2^50: centred pairs as for the
    // twiddles (mkf), appended to the same image (16-byte entries first)
    bool f64 = true;
    for (size_t i = 0; i < I; ++i) f64 = f64 && q[i] < (1ull << 50);
    for (size_t j = 0; j < J; ++j) f64 = f64 && p[j] < (1ull << 50);
    const size_t int_words = img.size() + (img.size() & 1);  // keep the pairs 16-byte aligned
    size_t off_inv_f = 0, off_m_f = 0, off_qmp_f = 0, off_pd = 0, off_pinv = 0;
    if (f64) {
        std::vector<ShoupF> pairs;
        for (size_t i = 0; i < I; ++i) pairs.push_back(mkf(Shoup{inv[i], inv_q[i]}, q[i]));
        for (size_t j = 0; j < J; ++j)
            for (size_t i = 0; i < I; ++i)
                pairs.push_back(mkf(Shoup{m[j * I + i], h_shoup_quotient(m[j * I + i], p[j])}, p[j]));
        for (size_t j = 0; j < J; ++j) pairs.push_back(mkf(Shoup{q_mod_p[j], h_shoup_quotient(q_mod_p[j], p[j])}, p[j]));
        std::vector<double> scal;
        for (size_t j = 0; j < J; ++j) scal.push_back((double)p[j]);
        for (size_t j = 0; j < J; ++j) scal.push_back(1.0 / (double)p[j]);
        // (the host pointers q, p, m, ... into img are not used past this point)
        img.resize(int_words + pairs.size() * 2 + scal.size());
        std::memcpy(img.data() + int_words, pairs.data(), pairs.size() * sizeof(ShoupF));
        std::memcpy(img.data() + int_words + pairs.size() * 2, scal.data(), scal.size() * sizeof(double));
        off_inv_f = int_words;
        off_m_f = off_inv_f + 2 * I;
        off_qmp_f = off_m_f + 2 * J * I;
        off_pd = int_words + pairs.size() * 2;
        off_pinv = off_pd + J;
    }
    ph_base_converter *c = new (std::nothrow) ph_base_converter();
    if (!c) return fail(PH_ERR_NO_MEMORY, "host allocation failed");
    c->in_count = in_count;
    c->out_count = out_count;
    int rc = upload(&c->d_mem, img.data(), img.size() * 8);
    if (rc != PH_OK) {
        ph_base_converter_free(c);
        return rc;
    }
    const uint64_t *base = (const uint64_t *)c->d_mem;
    size_t o = 0;
    auto next = [&](size_t words) {
        const uint64_t *ptr = base + o;
        o += words;
        return ptr;
    };
    c->dev.in_count = in_count;
    c->dev.out_count = out_count;
    c->dev.q = next(I);
    c->dev.q_r0 = next(I);
    c->dev.q_r1 = next(I);
    c->dev.inv = next(I);
    c->dev.inv_q = next(I);
    c->dev.qd = (const double *)next(I);
    c->dev.p = next(J);
    c->dev.p_r0 = next(J);
    c->dev.p_r1 = next(J);
    c->dev.m = next(J * I);
    c->dev.q_mod_p = next(J);
    {
        // the sum of k products y_i (Q/q_i mod p) stays below p * 2^64 while
        // k * max q_i <= 2^64: barrett_reduce128_wide's range
        uint64_t qmax = 0;
        for (size_t i = 0; i < I; ++i) qmax = std::max(qmax, ibase[i]);
        const u128 k = ((u128)1 << 64) / qmax;
        c->dev.lazy_terms = (uint32_t)std::min<u128>(k, (u128)I);
    }
    c->dev.f64 = f64 ? 1u : 0u;
    if (f64) {
        c->dev.inv_f = (const ShoupF *)(base + off_inv_f);
        c->dev.m_f = (const ShoupF *)(base + off_m_f);
        c->dev.qmp_f = (const ShoupF *)(base + off_qmp_f);
        c->dev.pd = (const double *)(base + off_pd);
        c->dev.pinv = (const double *)(base + off_pinv);
    }
    *out = c;
    return PH_OK;
}

static int base_convert(const ph_base_converter *c, const uint64_t *d_in, uint64_t *d_out, uint32_t log_n,
                        size_t count, bool exact, void *stream) {
    if (!c || ((!d_in || !d_out) && count)) return fail(PH_ERR_INVALID, "null argument");
    if (exact && c->out_count != 1) return fail(PH_ERR_INVALID, "exact conversion has one output modulus");
    if (log_n > 20) return fail(PH_ERR_UNSUPPORTED, "ring degree above 2^20");
    if (count == 0) return PH_OK;
    const uint64_t in_bytes = (uint64_t)count * c->in_count << (log_n + 3);
    const uint64_t out_bytes = (uint64_t)count * c->out_count << (log_n + 3);
    const uintptr_t a = (uintptr_t)d_in, b = (uintptr_t)d_out;
    if (a < b + out_bytes && b < a + in_bytes) return fail(PH_ERR_INVALID, "input and output overlap");
    hipError_t e = launch_base_convert(c->dev, d_in, d_out, log_n, count, exact, (hipStream_t)stream);
    return e == hipSuccess ? PH_OK : hip_fail(e, "base conversion kernel");
}

int ph_fast_convert_array_device(const ph_base_converter *c, const uint64_t *d_in, uint64_t *d_out, uint32_t log_n,
                                 size_t count, void *stream) {
    return base_convert(c, d_in, d_out, log_n, count, false, stream);
}

int ph_exact_convert_array_device(const ph_base_converter *c, const uint64_t *d_in, uint64_t *d_out,
                                  uint32_t log_n, size_t count, void *stream) {
    return base_convert(c, d_in, d_out, log_n, count, true, stream);
}

int ph_device_alloc(size_t bytes, void **d_ptr) {
    if (!d_ptr) return fail(PH_ERR_INVALID, "null output");
    PH_HIP(hipMalloc(d_ptr, bytes));
    return PH_OK;
}

int ph_device_free(void *d_ptr) {
    PH_HIP(hipFree(d_ptr));
    return PH_OK;
}

int ph_memcpy_htod(void *d_dst, const void *h_src, size_t bytes) {
    PH_HIP(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
    return PH_OK;
}

int ph_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes) {
    PH_HIP(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
    return PH_OK;
}

int ph_stream_synchronize(void *stream) {
    PH_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PH_OK;
}

int ph_host_alloc(size_t bytes, void **h_ptr) {
    if (!h_ptr) return fail(PH_ERR_INVALID, "null output");
    PH_HIP(hipHostMalloc(h_ptr, bytes, hipHostMallocDefault));
    return PH_OK;
}

int ph_host_free(void *h_ptr) {
    PH_HIP(hipHostFree(h_ptr));
    return PH_OK;
}

// ---------------------------------------------------------------------------
// external product
// ---------------------------------------------------------------------------

struct ph_ext_key {
    const ph_ntt_table *table;  // not owned
    ExtArgs proto;              // gadget constants, table and key pointers
    Shoup *d_key = nullptr;
    ShoupF *d_key_f = nullptr;
    bool f64 = false;
};

int ph_ext_key_new(const ph_ntt_table *t, uint32_t log_basis, uint32_t decompose_length, const uint64_t *key_ntt,
                   ph_ext_key **out) {
    return ph_ext_key_new_digits(t, log_basis, decompose_length, PH_DIGITS_APPROX_SIGNED, key_ntt, out);
}

int ph_ext_key_new_digits(const ph_ntt_table *t, uint32_t log_basis, uint32_t decompose_length, int digits,
                          const uint64_t *key_ntt, ph_ext_key **out) {
    if (!t || !key_ntt || !out) return fail(PH_ERR_INVALID, "null argument");
    *out = nullptr;
    if (digits != PH_DIGITS_APPROX_SIGNED && digits != PH_DIGITS_UNSIGNED)
        return fail(PH_ERR_INVALID, "digits: PH_DIGITS_APPROX_SIGNED or PH_DIGITS_UNSIGNED");
    if (t->log_n < kExtMinLogN || t->log_n > kExtMaxLogN)
        return fail(PH_ERR_UNSUPPORTED, "external product supports 2^6 <= N <= 2^12");
    const uint64_t q = t->q, n = t->n;
    // BigUintApproxSignedBasis::new(q, log_basis, Some(l)) for one limb
    // (primus_decompose/src/big_integer/basis.rs:34-195)
    const uint32_t bits = 64 - __builtin_clzll(q);
    if (log_basis == 0 || log_basis >= 32 || decompose_length == 0 || bits / log_basis < decompose_length)
        return fail(PH_ERR_INVALID, "gadget basis does not fit the modulus");
    ExtArgs g{};
    g.l = decompose_length;
    g.log_basis = log_basis;
    g.drop_bits = bits - decompose_length * log_basis;
    const uint64_t basis = 1ull << log_basis;
    g.basis_minus_one = basis - 1;
    g.carry_mask = log_basis == 1 ? 2ull : (basis | (basis >> 1));
    g.half = (basis + 1) / 2;
    g.basis_is_two = basis == 2;
    g.q_minus_basis = q - basis;
    u128 split = 0;
    bool has_split;
    if (log_basis == 1) {
        if (g.drop_bits == 0) {
            has_split = false;
        } else {
            for (uint32_t i = 0; i <= decompose_length; ++i) split = (split << 1) | 1;
            split <<= (g.drop_bits - 1);
            has_split = split < q;
        }
    } else {
        for (uint32_t i = 0; i < decompose_length; ++i) split = (split << log_basis) | (g.basis_minus_one >> 1);
        if (g.drop_bits > 0) split = ((split << 1) | 1) << (g.drop_bits - 1);
        else split += 1;
        has_split = split < q;
    }
    g.has_split = has_split;
    g.split = (uint64_t)split;
    g.pow2_sub_q = (uint64_t)((((u128)1) << bits) - q);
    g.tab = t->d_tab;
    g.mac_reduce_every = ext_mac_reduce_interval(q);
    if (digits == PH_DIGITS_UNSIGNED) {
        g.unsigned_digits = 1;
        g.has_split = 0;
    }

    // key rows [2][l][2][N] in NTT form, permuted so that slot t*E + j of a row
    // sits at j*(N/E) + t (the register layout of the kernel's last round)
    const uint64_t rows = 4ull * decompose_length, E = 1ull << kExtLogE, T = n / E;
    std::vector<Shoup> hk(rows * n);
    std::vector<ShoupF> hf;
    const bool f64 = t->f64;
    if (f64) hf.resize(rows * n);
    for (uint64_t r = 0; r < rows; ++r)
        for (uint64_t k = 0; k < n; ++k) {
            const uint64_t v = key_ntt[r * n + k];
            if (v >= q) return fail(PH_ERR_INVALID, "key values must be canonical (< q)");
            const uint64_t pos = r * n + (k % E) * T + k / E;
            hk[pos] = mk(v, q);
            if (f64) hf[pos] = mkf(hk[pos], q);
        }
    ph_ext_key *key = new (std::nothrow) ph_ext_key();
    if (!key) return fail(PH_ERR_NO_MEMORY, "host allocation failed");
    key->table = t;
    key->f64 = f64;
    int rc = upload((void **)&key->d_key, hk.data(), hk.size() * sizeof(Shoup));
    if (rc == PH_OK && f64) rc = upload((void **)&key->d_key_f, hf.data(), hf.size() * sizeof(ShoupF));
    if (rc != PH_OK) {
        ph_ext_key_free(key);
        return rc;
    }
    g.key = key->d_key;
    g.key_f = key->d_key_f;
    key->proto = g;
    *out = key;
    return PH_OK;
}

uint32_t ph_ext_key_log_n(const ph_ext_key *key) { return key ? key->table->log_n : 0; }

// ---------------------------------------------------------------------------
// External product over several limbs
// ---------------------------------------------------------------------------

namespace {
// Little-endian kExtRnsWords-word integers on the host, for the basis constants.
using Words = std::array<uint64_t, kExtRnsWords>;

Words words_mul_small(const Words &v, uint64_t m, bool *overflow) {
    Words r{};
    u128 carry = 0;
    for (uint32_t i = 0; i < kExtRnsWords; ++i) {
        const u128 t = (u128)v[i] * m + carry;
        r[i] = (uint64_t)t;
        carry = t >> 64;
    }
    if (overflow) *overflow = carry != 0;
    return r;
}

uint64_t words_mod_small(const Words &v, uint64_t q) {
    u128 r = 0;
    for (int i = kExtRnsWords - 1; i >= 0; --i) r = ((r << 64) | v[i]) % q;
    return (uint64_t)r;
}

uint32_t words_bits(const Words &v) {
    for (int i = kExtRnsWords - 1; i >= 0; --i)
        if (v[i]) return 64u * i + 64u - (uint32_t)__builtin_clzll(v[i]);
    return 0;
}

bool words_less(const Words &a, const Words &b) {
    for (int i = kExtRnsWords - 1; i >= 0; --i)
        if (a[i] != b[i]) return a[i] < b[i];
    return false;
}

// v << s for s < 64 (bits shifted out of the top word are lost)
Words words_shl(const Words &v, uint32_t s) {
    if (s == 0) return v;
    Words r{};
    for (int i = kExtRnsWords - 1; i >= 0; --i) r[i] = (v[i] << s) | (i ? v[i - 1] >> (64 - s) : 0);
    return r;
}

Words words_shl_any(Words v, uint32_t s) {
    while (s >= 63) {
        v = words_shl(v, 63);
        s -= 63;
    }
    return words_shl(v, s);
}
}  // namespace

struct ph_ext_key_rns {
    const ph_dcrt_table *table;  // not owned
    ExtRnsArgs proto;
    Shoup *d_key = nullptr;  // the key with Shoup quotients
    // the fused kernel's copies, rows permuted for its register layout
    // (slot t*E + j at j*(N/E) + t), when it applies
    bool fused = false;
    int arith = 0;
    ExtRnsGarner garner{};  // the fused kernel's composition constants
    Shoup *d_key_perm = nullptr;
    ShoupF *d_key_perm_f = nullptr;
};

int ph_ext_key_rns_new(const ph_dcrt_table *d, uint32_t log_basis, uint32_t decompose_length, const uint64_t *key_ntt,
                       ph_ext_key_rns **out) {
    if (!d || !key_ntt || !out) return fail(PH_ERR_INVALID, "null argument");
    *out = nullptr;
    if (d->L == 0 || d->L > kExtRnsMaxLimbs) return fail(PH_ERR_UNSUPPORTED, "1 to 4 limbs");
    ExtRnsArgs g{};
    g.L = d->L;
    g.log_n = d->log_n;
    // Q and the punctured products Q/q_i (primus_rns/src/base.rs:33-70)
    Words Q{};
    Q[0] = 1;
    for (uint32_t i = 0; i < d->L; ++i) Q = words_mul_small(Q, d->tables[i]->q, nullptr);
    for (uint32_t i = 0; i < d->L; ++i) {
        const uint64_t q = d->tables[i]->q;
        Words p{};
        p[0] = 1;
        for (uint32_t j = 0; j < d->L; ++j)
            if (j != i) p = words_mul_small(p, d->tables[j]->q, nullptr);
        const uint64_t inv = inv_mod(words_mod_small(p, q), q);
        if (inv == 0) return fail(PH_ERR_INVALID, "limb moduli are not coprime");
        g.q[i] = q;
        g.inv[i] = inv;
        g.inv_q[i] = h_shoup_quotient(inv, q);
        for (uint32_t w = 0; w < kExtRnsWords; ++w) {
            g.punct[i][w] = p[w];
            if (p[w]) g.punct_words = std::max(g.punct_words, w + 1);
        }
    }
    // BigUintApproxSignedBasis::new(Q, log_basis, Some(l)) (primus_decompose/src/big_integer/basis.rs:34-195)
    const uint32_t bits = words_bits(Q);
    if (log_basis == 0 || log_basis >= 64 || decompose_length == 0 || bits / log_basis < decompose_length)
        return fail(PH_ERR_INVALID, "gadget basis does not fit the modulus");
    // wrapping_decompose_small_values_inplace maps a digit d >= B/2 to
    // q_i - B + d, which needs every limb modulus above the basis
    // (primus_rns/src/base.rs:209-213)
    for (uint32_t i = 0; i < d->L; ++i)
        if (d->tables[i]->q <= (1ull << log_basis))
            return fail(PH_ERR_INVALID, "every limb modulus must exceed the gadget basis 2^log_basis");
    g.l = decompose_length;
    g.log_basis = log_basis;
    g.drop_bits = bits - decompose_length * log_basis;
    g.basis = 1ull << log_basis;
    g.basis_minus_one = g.basis - 1;
    g.carry_mask = log_basis == 1 ? 2ull : (g.basis | (g.basis >> 1));
    g.half = (g.basis + 1) / 2;
    Words split{};
    bool has_split;
    if (log_basis == 1) {
        if (g.drop_bits == 0) {
            has_split = false;
        } else {
            for (uint32_t i = 0; i <= decompose_length; ++i) {
                split = words_shl(split, 1);
                split[0] |= 1;
            }
            split = words_shl_any(split, g.drop_bits - 1);
            has_split = words_less(split, Q);
        }
    } else {
        for (uint32_t i = 0; i < decompose_length; ++i) {
            split = words_shl(split, log_basis);
            split[0] |= g.basis_minus_one >> 1;
        }
        if (g.drop_bits > 0) {
            split = words_shl(split, 1);
            split[0] |= 1;
            split = words_shl_any(split, g.drop_bits - 1);
        } else {
            u128 c = 1;  // split + 1
            for (uint32_t w = 0; w < kExtRnsWords; ++w) {
                const u128 t = (u128)split[w] + c;
                split[w] = (uint64_t)t;
                c = t >> 64;
            }
        }
        has_split = words_less(split, Q);
    }
    g.has_split = has_split;
    // 2^bits - Q: the two's complement of Q in kExtRnsWords words, reduced mod 2^bits
    Words pow2_sub{};
    u128 c = 1;
    for (uint32_t w = 0; w < kExtRnsWords; ++w) {
        const u128 t = (u128)(~Q[w]) + c;
        pow2_sub[w] = (uint64_t)t;
        c = t >> 64;
    }
    for (uint32_t w = 0; w < kExtRnsWords; ++w) {
        const uint32_t lo = 64 * w;
        if (bits <= lo) pow2_sub[w] = 0;
        else if (bits < lo + 64) pow2_sub[w] &= (1ull << (bits - lo)) - 1;
    }
    for (uint32_t w = 0; w < kExtRnsWords; ++w) {
        g.Q[w] = Q[w];
        g.split[w] = split[w];
        g.pow2_sub[w] = pow2_sub[w];
    }
    // key [2][l][2][L][N], canonical per limb
    const uint64_t n = 1ull << d->log_n, words = 4ull * decompose_length * d->L * n;
    for (uint64_t r = 0; r < 4ull * decompose_length * d->L; ++r) {
        const uint64_t q = d->tables[r % d->L]->q;
        for (uint64_t k = 0; k < n; ++k)
            if (key_ntt[r * n + k] >= q) return fail(PH_ERR_INVALID, "key values must be below their limb modulus");
    }
    std::vector<Shoup> hk(words);
    for (uint64_t r = 0; r < 4ull * decompose_length * d->L; ++r) {
        const uint64_t q = d->tables[r % d->L]->q;
        for (uint64_t k = 0; k < n; ++k) hk[r * n + k] = mk(key_ntt[r * n + k], q);
    }
    ph_ext_key_rns *key = new (std::nothrow) ph_ext_key_rns();
    if (!key) return fail(PH_ERR_NO_MEMORY, "key");
    key->table = d;
    if (hipMalloc(&key->d_key, words * sizeof(Shoup)) != hipSuccess ||
        hipMemcpy(key->d_key, hk.data(), words * sizeof(Shoup), hipMemcpyHostToDevice) != hipSuccess) {
        if (key->d_key) (void)hipFree(key->d_key);
        delete key;
        return fail(PH_ERR_DEVICE, "key upload");
    }
    // The fused kernel: single-pass ring degrees, digits within one 64-bit
    // window of the composed value (l * log_basis bits, plus the carry bit
    // below them when bits are dropped).
    key->fused = d->log_n >= kExtMinLogN && d->log_n <= kExtMaxLogN &&
                 (uint64_t)decompose_length * log_basis + (g.drop_bits > 0 ? 1 : 0) <= 64;
    if (key->fused) {
        key->arith = arith_of(d);
        uint64_t q_max = 0;
        for (uint32_t i = 0; i < d->L; ++i) q_max = std::max(q_max, d->tables[i]->q);
        g.mac_reduce_every = ext_mac_reduce_interval(q_max);
        // Garner's mixed-radix constants over the limbs in table order
        ExtRnsGarner &gk = key->garner;
        for (uint32_t i = 1; i < d->L; ++i) {
            const uint64_t qi = d->tables[i]->q;
            uint64_t prefix = 1;  // q_0 ... q_{i-1} mod q_i
            for (uint32_t j = 0; j < i; ++j) prefix = (uint64_t)((u128)prefix * (d->tables[j]->q % qi) % qi);
            gk.c[i] = inv_mod(prefix, qi);
            gk.c_q[i] = h_shoup_quotient(gk.c[i], qi);
            gk.off[i] = qi * (((1ull << 63) + qi - 1) / qi);
            for (uint32_t j = 0; j + 1 < i; ++j) {
                gk.qm[i][j] = d->tables[j]->q % qi;
                gk.qm_q[i][j] = h_shoup_quotient(gk.qm[i][j], qi);
            }
        }
        const uint64_t E = 1ull << kExtLogE, T = n / E, rows = 4ull * decompose_length * d->L;
        std::vector<Shoup> hp(words);
        std::vector<ShoupF> hpf;
        if (d->f64) hpf.resize(words);
        for (uint64_t r = 0; r < rows; ++r) {
            const uint64_t q = d->tables[r % d->L]->q;
            for (uint64_t k = 0; k < n; ++k) {
                const uint64_t pos = r * n + (k % E) * T + k / E;
                hp[pos] = hk[r * n + k];
                if (d->f64) hpf[pos] = mkf(hp[pos], q);
            }
        }
        int rc = upload((void **)&key->d_key_perm, hp.data(), hp.size() * sizeof(Shoup));
        if (rc == PH_OK && d->f64) rc = upload((void **)&key->d_key_perm_f, hpf.data(), hpf.size() * sizeof(ShoupF));
        if (rc != PH_OK) {
            ph_ext_key_rns_free(key);
            return rc;
        }
    }
    key->proto = g;
    *out = key;
    return PH_OK;
}

void ph_ext_key_rns_free(ph_ext_key_rns *key) {
    if (!key) return;
    if (key->d_key) (void)hipFree(key->d_key);
    if (key->d_key_perm) (void)hipFree(key->d_key_perm);
    if (key->d_key_perm_f) (void)hipFree(key->d_key_perm_f);
    delete key;
}

int ph_ext_product_rns_device(const ph_ext_key_rns *key, const uint64_t *d_in, uint64_t *d_out, size_t count,
                              int ntt_output, void *stream) {
    if (!key || ((!d_in || !d_out) && count)) return fail(PH_ERR_INVALID, "null argument");
    if (count == 0) return PH_OK;
    const ph_dcrt_table *d = key->table;
    hipStream_t s = (hipStream_t)stream;
    if (key->fused) {
        // one launch, nothing allocated: capturable in a HIP graph
        ExtRnsFusedArgs f{};
        f.g = key->proto;
        f.g.in = d_in;
        f.g.digits = nullptr;
        f.g.count = count;
        f.out = d_out;
        f.key = key->d_key_perm;
        f.key_f = key->d_key_perm_f;
        f.tabs = d->d_tabs;
        f.ntt_output = ntt_output ? 1u : 0u;
        f.garner = key->garner;
        hipError_t e = launch_ext_rns_fused(f, key->arith, s);
        return e == hipSuccess ? PH_OK : hip_fail(e, "fused external product kernel");
    }
    const uint64_t n = 1ull << d->log_n, L = d->L, l = key->proto.l;
    uint64_t *digits = nullptr;
    PH_HIP(hipMallocAsync((void **)&digits, count * 2 * l * L * n * 8, s));
    ExtRnsArgs a = key->proto;
    a.in = d_in;
    a.digits = digits;
    a.count = count;
    hipError_t e = launch_ext_rns_digits(a, s);
    int rc = e == hipSuccess ? PH_OK : hip_fail(e, "digit kernel");
    // the digits' forward transforms: row r of [count][2][l][L][N] is limb r % L
    if (rc == PH_OK) rc = run_op(Op::FwdCanon, d->d_tabs, d->L, d->log_n, arith_of(d), digits, nullptr, nullptr,
                                 count * 2 * l * L, s);
    if (rc == PH_OK) {
        ExtRnsMacArgs m{};
        m.digits = digits;
        m.key = key->d_key;
        m.out = d_out;
        m.tabs = d->d_tabs;
        m.count = count;
        m.L = d->L;
        m.l = (uint32_t)l;
        m.log_n = d->log_n;
        e = launch_ext_rns_mac(m, s);
        if (e != hipSuccess) rc = hip_fail(e, "multiply-accumulate kernel");
    }
    if (rc == PH_OK && !ntt_output)
        rc = run_op(Op::Inv, d->d_tabs, d->L, d->log_n, arith_of(d), d_out, nullptr, nullptr, count * 2 * L, s);
    (void)hipFreeAsync(digits, s);
    return rc;
}

void ph_ext_key_free(ph_ext_key *key) {
    if (!key) return;
    if (key->d_key) (void)hipFree(key->d_key);
    if (key->d_key_f) (void)hipFree(key->d_key_f);
    delete key;
}

int ph_ext_product_device(const ph_ext_key *key, const uint64_t *d_in, uint64_t *d_out, size_t count, int ntt_output,
                          void *stream) {
    if (!key || ((!d_in || !d_out) && count)) return fail(PH_ERR_INVALID, "null argument");
    ExtArgs g = key->proto;
    g.in = d_in;
    g.out = d_out;
    g.ntt_output = ntt_output ? 1u : 0u;
    hipError_t e = launch_ext_product(g, key->table->log_n, key->f64, count, (hipStream_t)stream);
    return e == hipSuccess ? PH_OK : hip_fail(e, "external product kernel");
}

int ph_ext_product(const ph_ext_key *key, const uint64_t *in, uint64_t *out, size_t count, int ntt_output) {
    if (!key || ((!in || !out) && count)) return fail(PH_ERR_INVALID, "null argument");
    if (count == 0) return PH_OK;
    const size_t words = count * 2 * key->table->n;
    void *buf;
    hipStream_t s;
    int rc = staging(2 * words * 8, &buf, &s);
    if (rc != PH_OK) return rc;
    uint64_t *d_in = (uint64_t *)buf, *d_out = d_in + words;
    PH_HIP(hipMemcpyAsync(d_in, in, words * 8, hipMemcpyHostToDevice, s));
    rc = ph_ext_product_device(key, d_in, d_out, count, ntt_output, s);
    if (rc != PH_OK) return rc;
    PH_HIP(hipMemcpyAsync(out, d_out, words * 8, hipMemcpyDeviceToHost, s));
    PH_HIP(hipStreamSynchronize(s));
    return PH_OK;
}

// ---------------------------------------------------------------------------
// RLWE encryption
// ---------------------------------------------------------------------------

struct ph_rlwe_ctx {
    uint32_t min_log_n, max_log_n, max_limbs;
    std::vector<ph_ntt_table *> tables;  // [(log_n - min_log_n) * max_limbs + i]
    DevTable *d_tabs = nullptr;
    int cus = 0;
    bool wide = true;  // every modulus below 2^60: IntCanonArith kernels
};

void rlwe_ctx_free(ph_rlwe_ctx *c) {
    if (!c) return;
    for (ph_ntt_table *t : c->tables) free_table(t);
    if (c->d_tabs) (void)hipFree(c->d_tabs);
    delete c;
}

int ph_rlwe_ctx_new(uint32_t min_log_n, uint32_t max_log_n, uint32_t max_limbs, const uint64_t *moduli,
                    ph_rlwe_ctx **out) {
    if (!out || !moduli || max_limbs == 0 || min_log_n > max_log_n) return fail(PH_ERR_INVALID, "bad arguments");
    *out = nullptr;
    if (min_log_n < kRlweMinLogN || max_log_n > kRlweMaxLogN)
        return fail(PH_ERR_UNSUPPORTED, "RLWE encryption supports 2^10 <= N <= 2^16");
    ph_rlwe_ctx *c = new (std::nothrow) ph_rlwe_ctx();
    if (!c) return fail(PH_ERR_NO_MEMORY, "host allocation failed");
    c->min_log_n = min_log_n;
    c->max_log_n = max_log_n;
    c->max_limbs = max_limbs;
    std::vector<DevTable> views;
    for (uint32_t ln = min_log_n; ln <= max_log_n; ++ln)
        for (uint32_t i = 0; i < max_limbs; ++i) {
            ph_ntt_table *t = nullptr;
            int rc = build_table(ln, moduli[(ln - min_log_n) * max_limbs + i], PH_TABLE_INTEGER_ONLY, &t);
            if (rc != PH_OK) {
                rlwe_ctx_free(c);
                return rc;
            }
            c->tables.push_back(t);
            c->wide = c->wide && t->wide;
            views.push_back(t->dev);
        }
    int rc = upload((void **)&c->d_tabs, views.data(), views.size() * sizeof(DevTable));
    if (rc == PH_OK && hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess)
        rc = fail(PH_ERR_DEVICE, "cannot query the compute-unit count");
    if (rc != PH_OK) {
        rlwe_ctx_free(c);
        return rc;
    }
    *out = c;
    return PH_OK;
}

void ph_rlwe_ctx_free(ph_rlwe_ctx *c) { rlwe_ctx_free(c); }

uint64_t ph_rlwe_ctx_modulus(const ph_rlwe_ctx *c, uint32_t log_n, uint32_t limb) {
    if (!c || log_n < c->min_log_n || log_n > c->max_log_n || limb >= c->max_limbs) return 0;
    return c->tables[(log_n - c->min_log_n) * c->max_limbs + limb]->q;
}

// One work item per (request, limb), largest ring degrees first (the order
// the kernel's work queue hands them out), with Delta mod q_i, and one zeroed
// element after them: the launch's work queue (RlweArgs::queue).
static int rlwe_items(const ph_rlwe_ctx *c, const ph_rlwe_request *reqs, uint32_t n_reqs,
                      std::vector<RlweItem> &items) {
    items.clear();
    for (uint32_t r = 0; r < n_reqs; ++r) {
        const ph_rlwe_request &q = reqs[r];
        if (q.log_n < c->min_log_n || q.log_n > c->max_log_n || q.limbs == 0 || q.limbs > c->max_limbs)
            return fail(PH_ERR_INVALID, "request outside the context's ring degrees or limb count");
        const uint64_t n = 1ull << q.log_n;
        const uint32_t base = (q.log_n - c->min_log_n) * c->max_limbs;
        // Delta = floor(Q / 256), Q the product of the request's moduli, kept
        // per limb as Delta mod q_i = (-(Q mod 256)) * 256^-1 mod q_i.
        uint64_t q_mod_256 = 1;
        for (uint32_t i = 0; i < q.limbs; ++i) q_mod_256 = (q_mod_256 * (c->tables[base + i]->q & 255)) & 255;
        for (uint32_t i = 0; i < q.limbs; ++i) {
            const uint64_t qi = c->tables[base + i]->q;
            const uint64_t inv256 = inv_mod(256 % qi, qi);
            const uint64_t delta = h_mulmod((qi - q_mod_256 % qi) % qi, inv256, qi);
            RlweItem it{};
            it.row = q.a_offset + (uint64_t)i * n;
            it.small = q.small_offset;
            it.delta = delta;
            it.delta_q = h_shoup_quotient(delta, qi);
            it.log_n = q.log_n;
            it.tab = base + i;
            items.push_back(it);
        }
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const RlweItem &x, const RlweItem &y) { return x.log_n > y.log_n; });
    items.push_back(RlweItem{});
    return PH_OK;
}

// Persistent grid of an RLWE launch over n_items items: two resident
// workgroups per compute unit (the kernel's occupancy), fewer for a short list.
static uint32_t rlwe_grid(const ph_rlwe_ctx *c, uint64_t n_items) {
    return (uint32_t)std::min<uint64_t>(n_items, (uint64_t)std::max(c->cus, 1) * 2);
}

// Scratch words per workgroup: only the two-pass ring degrees use scratch,
// one polynomial of the largest ring degree.
static uint64_t rlwe_scratch_words(uint32_t max_log_n) { return max_log_n > 12 ? 1ull << max_log_n : 0; }

static hipError_t rlwe_launch(const ph_rlwe_ctx *c, RlweItem *d_items, uint64_t n_items, uint32_t max_log_n,
                              const uint64_t *d_a, const int8_t *d_s, const int8_t *d_e, const uint8_t *d_m,
                              uint64_t *d_out, uint64_t *scratch, hipStream_t s) {
    RlweArgs a{};
    a.items = d_items;
    a.n_items = (uint32_t)n_items;
    a.queue = (uint32_t *)(d_items + n_items);
    a.tabs = c->d_tabs;
    a.a = d_a;
    a.out = d_out;
    a.s = d_s;
    a.e = d_e;
    a.m = d_m;
    a.scratch = scratch;
    a.scratch_words = rlwe_scratch_words(max_log_n);
    a.canon_int = c->wide ? 1u : 0u;
    return launch_rlwe_encrypt(a, rlwe_grid(c, n_items), s);
}

int ph_rlwe_encrypt_device(const ph_rlwe_ctx *c, const ph_rlwe_request *reqs, uint32_t n_reqs, const uint64_t *d_a,
                           const int8_t *d_s, const int8_t *d_e, const uint8_t *d_m, uint64_t *d_out, void *stream) {
    if (!c || (!reqs && n_reqs)) return fail(PH_ERR_INVALID, "null argument");
    if (n_reqs == 0) return PH_OK;
    if (!d_a || !d_s || !d_e || !d_m || !d_out) return fail(PH_ERR_INVALID, "null device buffer");
    hipStream_t s = (hipStream_t)stream;
    std::vector<RlweItem> items;
    int rc = rlwe_items(c, reqs, n_reqs, items);
    if (rc != PH_OK) return rc;
    RlweItem *d_items = nullptr;
    uint64_t *scratch = nullptr;
    const uint64_t n_items = items.size() - 1;  // and the queue
    PH_HIP(hipMallocAsync((void **)&d_items, items.size() * sizeof(RlweItem), s));
    PH_HIP(hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(RlweItem), hipMemcpyHostToDevice, s));
    const uint32_t max_log_n = items.front().log_n;
    const uint64_t scratch_words = rlwe_scratch_words(max_log_n);
    if (scratch_words)
        PH_HIP(hipMallocAsync((void **)&scratch, (size_t)rlwe_grid(c, n_items) * scratch_words * 8, s));
    PH_HIP(hipStreamSynchronize(s));  // the host item list may go away after this
    hipError_t e = rlwe_launch(c, d_items, n_items, max_log_n, d_a, d_s, d_e, d_m, d_out, scratch, s);
    if (scratch) (void)hipFreeAsync(scratch, s);
    (void)hipFreeAsync(d_items, s);
    return e == hipSuccess ? PH_OK : hip_fail(e, "RLWE encryption kernel");
}

// Cumulative thresholds of the rounded Gaussian (sigma 3.2, |e| <= 19): the
// weights exp(-k^2 / (2 sigma^2)) for k = -19 .. 19 summed in order, and
// threshold i = floor(2^64 * (w_-19 + ... + w_(-19+i)) / sum).
static void gauss_thresholds(uint64_t *t) {
    const double sigma = 3.2;
    double w[2 * kGaussCut + 1], z = 0.0;
    for (int k = -kGaussCut; k <= kGaussCut; ++k) {
        w[k + kGaussCut] = std::exp(-(double)(k * k) / (2.0 * sigma * sigma));
        z += w[k + kGaussCut];
    }
    double cum = 0.0;
    for (int i = 0; i < kGaussThresholds; ++i) {
        cum += w[i];
        t[i] = (uint64_t)std::ldexp(cum / z, 64);
    }
}

// The sampling launch's chunks (SampleChunk) for a request list; the
// cumulative Gaussian thresholds follow them in the same upload.
static int sample_chunks(const ph_rlwe_ctx *c, const ph_rlwe_request *reqs, uint32_t n_reqs,
                         std::vector<SampleChunk> &chunks) {
    chunks.clear();
    for (uint32_t r = 0; r < n_reqs; ++r) {
        const ph_rlwe_request &q = reqs[r];
        if (q.log_n < c->min_log_n || q.log_n > c->max_log_n || q.limbs == 0 || q.limbs > c->max_limbs)
            return fail(PH_ERR_INVALID, "request outside the context's ring degrees or limb count");
        const uint32_t n = 1u << q.log_n;
        const uint32_t base = (q.log_n - c->min_log_n) * c->max_limbs;
        for (uint32_t k0 = 0; k0 < n; k0 += kSampleChunk) {
            const uint32_t cnt = std::min(kSampleChunk, n - k0);
            for (uint32_t p : {(uint32_t)kSampleSecret, (uint32_t)kSampleError, (uint32_t)kSampleMessage})
                chunks.push_back(SampleChunk{q.small_offset + k0, 0, r, p, k0, cnt});
            for (uint32_t j = 0; j < q.limbs; ++j)
                chunks.push_back(SampleChunk{q.a_offset + (uint64_t)j * n + k0, c->tables[base + j]->q, r,
                                             (uint32_t)kSampleA + j, k0, cnt});
        }
    }
    return PH_OK;
}

static hipError_t sample_launch(const char *d_mem, uint32_t n_chunks, uint64_t seed, uint64_t *d_a, int8_t *d_s,
                                int8_t *d_e, uint8_t *d_m, hipStream_t s) {
    SampleArgs a{};
    a.chunks = (const SampleChunk *)d_mem;
    a.n_chunks = n_chunks;
    a.seed_lo = (uint32_t)seed;
    a.seed_hi = (uint32_t)(seed >> 32);
    a.gauss = (const uint64_t *)(d_mem + (size_t)n_chunks * sizeof(SampleChunk));
    a.a = d_a;
    a.s = d_s;
    a.e = d_e;
    a.m = d_m;
    return launch_rlwe_sample(a, s);
}

int ph_rlwe_sample_device(const ph_rlwe_ctx *c, const ph_rlwe_request *reqs, uint32_t n_reqs, uint64_t seed,
                          uint64_t *d_a, int8_t *d_s, int8_t *d_e, uint8_t *d_m, void *stream) {
    if (!c || (!reqs && n_reqs)) return fail(PH_ERR_INVALID, "null argument");
    if (n_reqs == 0) return PH_OK;
    if (!d_a || !d_s || !d_e || !d_m) return fail(PH_ERR_INVALID, "null device buffer");
    std::vector<SampleChunk> chunks;
    int rc = sample_chunks(c, reqs, n_reqs, chunks);
    if (rc != PH_OK) return rc;
    uint64_t gauss[kGaussThresholds];
    gauss_thresholds(gauss);
    hipStream_t s = (hipStream_t)stream;
    const size_t cb = chunks.size() * sizeof(SampleChunk);
    char *d_mem = nullptr;
    PH_HIP(hipMallocAsync((void **)&d_mem, cb + sizeof(gauss), s));
    PH_HIP(hipMemcpyAsync(d_mem, chunks.data(), cb, hipMemcpyHostToDevice, s));
    PH_HIP(hipMemcpyAsync(d_mem + cb, gauss, sizeof(gauss), hipMemcpyHostToDevice, s));
    PH_HIP(hipStreamSynchronize(s));  // the host chunk list may go away after this
    hipError_t e = sample_launch(d_mem, (uint32_t)chunks.size(), seed, d_a, d_s, d_e, d_m, s);
    (void)hipFreeAsync(d_mem, s);
    return e == hipSuccess ? PH_OK : hip_fail(e, "RLWE sampling kernel");
}

struct ph_rlwe_plan {
    const ph_rlwe_ctx *ctx = nullptr;
    RlweItem *d_items = nullptr;
    uint32_t n_items = 0;
    uint32_t grid = 0, max_grid = 0;
    uint64_t *scratch = nullptr;
    uint64_t scratch_words = 0;
    char *d_sample = nullptr;  // the sampling launch's chunks, then the Gaussian thresholds
    uint32_t n_chunks = 0;
};

void ph_rlwe_plan_free(ph_rlwe_plan *p) {
    if (!p) return;
    if (p->d_items) (void)hipFree(p->d_items);
    if (p->scratch) (void)hipFree(p->scratch);
    if (p->d_sample) (void)hipFree(p->d_sample);
    delete p;
}

int ph_rlwe_plan_new(const ph_rlwe_ctx *c, const ph_rlwe_request *reqs, uint32_t n_reqs, ph_rlwe_plan **out) {
    if (!c || !out || !reqs || n_reqs == 0) return fail(PH_ERR_INVALID, "bad arguments");
    *out = nullptr;
    std::vector<RlweItem> items;
    int rc = rlwe_items(c, reqs, n_reqs, items);
    if (rc != PH_OK) return rc;
    ph_rlwe_plan *p = new (std::nothrow) ph_rlwe_plan();
    if (!p) return fail(PH_ERR_NO_MEMORY, "host allocation failed");
    p->ctx = c;
    p->n_items = (uint32_t)items.size() - 1;  // the list's last element is the queue
    // room for up to four workgroups per compute unit (the autotuner's range);
    // two, the kernel's occupancy, by default
    const uint64_t cus = (uint64_t)std::max(c->cus, 1);
    p->max_grid = (uint32_t)std::min<uint64_t>(p->n_items, 4 * cus);
    p->grid = (uint32_t)std::min<uint64_t>(p->n_items, 2 * cus);
    p->scratch_words = items.front().log_n > 12 ? 1ull << items.front().log_n : 0;
    rc = upload((void **)&p->d_items, items.data(), items.size() * sizeof(RlweItem));
    if (rc == PH_OK && p->scratch_words && hipMalloc((void **)&p->scratch, (size_t)p->max_grid * p->scratch_words * 8) !=
                                               hipSuccess)
        rc = fail(PH_ERR_NO_MEMORY, "RLWE scratch");
    std::vector<SampleChunk> chunks;
    if (rc == PH_OK) rc = sample_chunks(c, reqs, n_reqs, chunks);
    if (rc == PH_OK) {
        std::vector<char> img(chunks.size() * sizeof(SampleChunk) + kGaussThresholds * sizeof(uint64_t));
        std::memcpy(img.data(), chunks.data(), chunks.size() * sizeof(SampleChunk));
        gauss_thresholds((uint64_t *)(img.data() + chunks.size() * sizeof(SampleChunk)));
        p->n_chunks = (uint32_t)chunks.size();
        rc = upload((void **)&p->d_sample, img.data(), img.size());
    }
    if (rc != PH_OK) {
        ph_rlwe_plan_free(p);
        return rc;
    }
    *out = p;
    return PH_OK;
}

int ph_rlwe_sample_plan_device(const ph_rlwe_plan *p, uint64_t seed, uint64_t *d_a, int8_t *d_s, int8_t *d_e,
                               uint8_t *d_m, void *stream) {
    if (!p) return fail(PH_ERR_INVALID, "null plan");
    if (!d_a || !d_s || !d_e || !d_m) return fail(PH_ERR_INVALID, "null device buffer");
    hipError_t e = sample_launch(p->d_sample, p->n_chunks, seed, d_a, d_s, d_e, d_m, (hipStream_t)stream);
    return e == hipSuccess ? PH_OK : hip_fail(e, "RLWE sampling kernel");
}

int ph_rlwe_encrypt_plan_device(const ph_rlwe_plan *p, const uint64_t *d_a, const int8_t *d_s, const int8_t *d_e,
                                const uint8_t *d_m, uint64_t *d_out, void *stream) {
    if (!p) return fail(PH_ERR_INVALID, "null plan");
    if (!d_a || !d_s || !d_e || !d_m || !d_out) return fail(PH_ERR_INVALID, "null device buffer");
    RlweArgs a{};
    a.items = p->d_items;
    a.n_items = p->n_items;
    a.queue = (uint32_t *)(p->d_items + p->n_items);
    a.tabs = p->ctx->d_tabs;
    a.a = d_a;
    a.out = d_out;
    a.s = d_s;
    a.e = d_e;
    a.m = d_m;
    a.scratch = p->scratch;
    a.scratch_words = p->scratch_words;
    a.canon_int = p->ctx->wide ? 1u : 0u;
    hipError_t e = launch_rlwe_encrypt(a, p->grid, (hipStream_t)stream);
    return e == hipSuccess ? PH_OK : hip_fail(e, "RLWE encryption kernel");
}

int ph_rlwe_plan_autotune(ph_rlwe_plan *p, const uint64_t *d_a, const int8_t *d_s, const int8_t *d_e,
                          const uint8_t *d_m, uint64_t *d_out, int reps, uint32_t *grid, float *best_ms) {
    if (!p || reps < 1) return fail(PH_ERR_INVALID, "bad autotune request");
    const uint64_t cus = (uint64_t)std::max(p->ctx->cus, 1);
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PH_OK;
    float best = 1e30f;
    uint32_t best_grid = p->grid;
    if (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fail(PH_ERR_DEVICE, "autotune set-up failed");
    for (uint64_t per_cu = 1; per_cu <= 4 && rc == PH_OK; ++per_cu) {
        const uint32_t g = (uint32_t)std::min<uint64_t>(p->n_items, per_cu * cus);
        p->grid = g;
        for (int r = 0; r <= reps && rc == PH_OK; ++r) {
            if (r == 1) (void)hipEventRecord(e0, s);
            rc = ph_rlwe_encrypt_plan_device(p, d_a, d_s, d_e, d_m, d_out, s);
        }
        if (rc != PH_OK) break;
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(PH_ERR_DEVICE, "autotune timing failed");
            break;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= reps;
        if (ms < best) {
            best = ms;
            best_grid = g;
        }
        if (g == p->n_items) break;  // larger grids would be the same launch
    }
    p->grid = best_grid;
    if (grid) *grid = best_grid;
    if (best_ms) *best_ms = best;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}

uint32_t ph_rlwe_plan_grid(const ph_rlwe_plan *p) { return p ? p->grid : 0; }

static_assert(PH_SHAPE_STREAM == kShapeStream && PH_SHAPE_LPB(1) == (1 << kShapeLpbShift) &&
                  PH_SHAPE_TILE(1) == (1 << kShapeTileShift),
              "the header's shape fields are the launcher's");
static_assert(PH_CLOCK_PROBE_SLOTS == kClockProbeSlots && PH_CLOCK_PROBE_DEPTH == kClockProbeDepth,
              "the header's probe log shape is the kernels'");

int ph_tuning_set(uint32_t log_n, int op, int arith, int shape) {
    if (log_n > kMaxLogN || op < 0 || op > 4 || arith < 0 || arith > 1 || !valid_shape_for(shape, log_n))
        return fail(PH_ERR_INVALID, "bad tuning key or value");
    g_tuning[log_n][op][arith].store(shape, std::memory_order_relaxed);
    return PH_OK;
}

int ph_tuning_shapes(uint32_t log_n, int op, int *shapes, int max) {
    if (log_n < kMinLogN || log_n > kMaxLogN || op < 0 || op > 4 || max < 0 || (max > 0 && !shapes))
        return -fail(PH_ERR_INVALID, "bad tuning-shape query");
    int tmp[64];
    const int n = tuning_candidates(log_n, (Op)op, tmp, 64);
    for (int i = 0; i < n && i < max; ++i) shapes[i] = tmp[i];
    return n;
}

int ph_tuning_get(uint32_t log_n, int op, int arith) {
    if (log_n > kMaxLogN || op < 0 || op > 4 || arith < 0 || arith > 1) return -1;
    return g_tuning[log_n][op][arith].load(std::memory_order_relaxed);
}

int ph_tuning_save(const char *path) {
    if (!path) return fail(PH_ERR_INVALID, "null path");
    FILE *f = fopen(path, "w");
    if (!f) return fail(PH_ERR_INVALID, "cannot open tuning file for writing");
    fprintf(f, "# primus_hip tuning: log_n op arith shape\n");
    for (uint32_t n = 0; n <= kMaxLogN; ++n)
        for (int op = 0; op < 5; ++op)
            for (int ar = 0; ar < 2; ++ar) {
                int v = g_tuning[n][op][ar].load();
                if (v) fprintf(f, "%u %d %d %d\n", n, op, ar, v);
            }
    fclose(f);
    return PH_OK;
}

int ph_tuning_load(const char *path) {
    if (!path) return fail(PH_ERR_INVALID, "null path");
    FILE *f = fopen(path, "r");
    if (!f) return fail(PH_ERR_INVALID, "cannot open tuning file");
    char line[256];
    int rc = PH_OK;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        unsigned n;
        int op, ar, v;
        if (sscanf(line, "%u %d %d %d", &n, &op, &ar, &v) == 4) {
            if (ph_tuning_set(n, op, ar, v) != PH_OK) rc = PH_ERR_INVALID;
        }
    }
    fclose(f);
    return rc;
}

int ph_autotune(const ph_dcrt_table *d, int op, size_t count, int reps, int *chosen, float *best_ms) {
    if (!d || op < 0 || op > 4 || count == 0 || reps < 1) return fail(PH_ERR_INVALID, "bad autotune request");
    const Op o = (Op)op;
    const bool f64 = d->f64 && o != Op::FwdLazy;
    const int arith = arith_of(d);
    const uint64_t n = 1ull << d->log_n;
    const size_t words = count * n;
    // Rotate over enough copies of the batch (up to 16) that the working set
    // exceeds 512 MiB: a cache-resident batch would favour other shapes than
    // a batch streamed from HBM.
    const size_t copies = std::min<size_t>(16, std::max<size_t>(1, ((512ull << 20) + words * 8 - 1) / (words * 8)));
    std::vector<uint64_t *> bufs(copies, nullptr), bufs2(copies, nullptr);
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PH_OK;
    std::vector<Shoup> ones;
    std::vector<ShoupF> onesv;
    ph_operand op_dev{};
    float best = 1e30f;
    int best_shape = 0;
    // the shape in force before the trials: put back if they fail part-way
    const int prior_shape = g_tuning[d->log_n][op][f64 ? 1 : 0].load();
    // every shape of this degree and operation (tuning_candidates): radix of
    // the register rounds, one workgroup per polynomial or the streaming kernel
    // with its limb-polynomials per workgroup, the two-pass LDS tile
    int candidates[64];
    const int ncand = tuning_candidates(d->log_n, o, candidates, 64);
    if (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = fail(PH_ERR_DEVICE, "autotune set-up failed");
        goto done;
    }
    for (size_t c = 0; c < copies; ++c) {
        const bool second = o == Op::PolyMul;
        if (hipMalloc(&bufs[c], words * 8) != hipSuccess || hipMemset(bufs[c], 0, words * 8) != hipSuccess ||
            (second && (hipMalloc(&bufs2[c], words * 8) != hipSuccess ||
                        hipMemset(bufs2[c], 0, words * 8) != hipSuccess))) {
            rc = fail(PH_ERR_NO_MEMORY, "autotune buffers");
            goto done;
        }
    }
    if (o == Op::FwdMulInv) {
        ones.assign(d->L * n, Shoup{1, 0});
        op_dev.L = d->L;
        op_dev.n = n;
        for (uint32_t l = 0; l < d->L; ++l)
            for (uint64_t i = 0; i < n; ++i) ones[l * n + i] = mk(1, d->tables[l]->q);
        if (upload((void **)&op_dev.d, ones.data(), ones.size() * sizeof(Shoup)) != PH_OK) {
            rc = PH_ERR_DEVICE;
            goto done;
        }
        if (d->f64) {
            onesv.resize(ones.size());
            for (uint32_t l = 0; l < d->L; ++l)  // stored times N^-1, as ph_operand_new does
                for (uint64_t i = 0; i < n; ++i)
                    onesv[l * n + i] = mkf(mk(d->tables[l]->inv_n, d->tables[l]->q), d->tables[l]->q);
            // all ones: every register order is the same array
            if (upload((void **)&op_dev.d_f[0], onesv.data(), onesv.size() * sizeof(ShoupF)) != PH_OK ||
                upload((void **)&op_dev.d_f[1], onesv.data(), onesv.size() * sizeof(ShoupF)) != PH_OK) {
                rc = PH_ERR_DEVICE;
                goto done;
            }
        }
    }
    for (int ci = 0; ci < ncand; ++ci) {
        const int shape = candidates[ci];
        g_tuning[d->log_n][op][f64 ? 1 : 0].store(shape);
        // warm-up, then timed repetitions
        for (int r = 0; r <= reps; ++r) {
            if (r == 1) (void)hipEventRecord(e0, s);
            rc = run_op(o, d->d_tabs, d->L, d->log_n, arith, bufs[r % copies], bufs2[r % copies],
                        o == Op::FwdMulInv ? &op_dev : nullptr, count, s);
            if (rc != PH_OK) goto done;
        }
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(PH_ERR_DEVICE, "autotune timing failed");
            goto done;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= reps;
        if (ms < best) {
            best = ms;
            best_shape = shape;
        }
    }
    g_tuning[d->log_n][op][f64 ? 1 : 0].store(best_shape);
    if (chosen) *chosen = best_shape;
    if (best_ms) *best_ms = best;
done:
    if (rc != PH_OK) g_tuning[d->log_n][op][f64 ? 1 : 0].store(prior_shape);
    if (op_dev.d) (void)hipFree(op_dev.d);
    for (ShoupF *p : op_dev.d_f)
        if (p) (void)hipFree(p);
    for (size_t c = 0; c < copies; ++c) {
        if (bufs[c]) (void)hipFree(bufs[c]);
        if (bufs2[c]) (void)hipFree(bufs2[c]);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}

}  // extern "C"

// Library-internal forms (capi_internal.hpp), C++ linkage.
namespace ph {

RlweStageSize rlwe_stage_size(const ph_rlwe_ctx *c, const ph_rlwe_request *reqs, uint32_t n_reqs) {
    RlweStageSize z{0, 0};
    if (!c || (!reqs && n_reqs)) return z;
    uint64_t n_items = 0;
    uint32_t max_log_n = 0;
    for (uint32_t r = 0; r < n_reqs; ++r) {
        const ph_rlwe_request &q = reqs[r];
        if (q.log_n < c->min_log_n || q.log_n > c->max_log_n || q.limbs == 0 || q.limbs > c->max_limbs) return z;
        n_items += q.limbs;
        max_log_n = std::max(max_log_n, q.log_n);
    }
    z.item_bytes = (n_items + 1) * sizeof(RlweItem);  // and the work queue
    z.scratch_bytes = rlwe_grid(c, n_items) * rlwe_scratch_words(max_log_n) * 8;
    return z;
}

int rlwe_encrypt_staged(const ph_rlwe_ctx *c, const ph_rlwe_request *reqs, uint32_t n_reqs, const uint64_t *d_a,
                        const int8_t *d_s, const int8_t *d_e, const uint8_t *d_m, uint64_t *d_out, hipStream_t s,
                        void *host_items, void *d_items, void *d_scratch, const RlweStageSize &capacity) {
    if (!c || (!reqs && n_reqs)) return fail(PH_ERR_INVALID, "null argument");
    if (n_reqs == 0) return PH_OK;
    if (!d_a || !d_s || !d_e || !d_m || !d_out || !host_items || !d_items) return fail(PH_ERR_INVALID, "null buffer");
    std::vector<RlweItem> items;
    int rc = rlwe_items(c, reqs, n_reqs, items);
    if (rc != PH_OK) return rc;
    const RlweStageSize need = rlwe_stage_size(c, reqs, n_reqs);
    if (need.item_bytes > capacity.item_bytes || need.scratch_bytes > capacity.scratch_bytes ||
        (need.scratch_bytes && !d_scratch))
        return fail(PH_ERR_INVALID, "staging buffers smaller than the request list needs");
    std::memcpy(host_items, items.data(), need.item_bytes);
    PH_HIP(hipMemcpyAsync(d_items, host_items, need.item_bytes, hipMemcpyHostToDevice, s));
    hipError_t e = rlwe_launch(c, (RlweItem *)d_items, items.size() - 1, items.front().log_n, d_a, d_s, d_e, d_m,
                               d_out, need.scratch_bytes ? (uint64_t *)d_scratch : nullptr, s);
    return e == hipSuccess ? PH_OK : hip_fail(e, "RLWE encryption kernel");
}

uint64_t negacyclic_mul_scratch_bytes(const ph_dcrt_table *d, uint64_t count) {
    return d && needs_scratch(Op::PolyMul, d->log_n) ? (count << d->log_n) * 8 : 0;
}

int negacyclic_mul_rows_staged(const ph_dcrt_table *d, uint64_t *d_a, const uint64_t *d_b, const uint8_t *d_limb_of,
                               uint64_t count, void *d_scratch, hipStream_t s) {
    if (!d || ((!d_a || !d_b || !d_limb_of) && count)) return fail(PH_ERR_INVALID, "null argument");
    if (negacyclic_mul_scratch_bytes(d, count) && !d_scratch) return fail(PH_ERR_INVALID, "null scratch");
    return run_op(Op::PolyMul, d->d_tabs, d->L, d->log_n, arith_of(d), d_a, d_b, nullptr, count, s, d_limb_of, d_scratch);
}

}  // namespace ph
