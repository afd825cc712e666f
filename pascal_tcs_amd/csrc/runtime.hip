// runtime.hip -- context lifecycle, module-array residency, halo, post
// reductions and the time-step composition behind the C-ABI of pascal_tcs.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include <algorithm>

#include "ad.h"
#include "tcs_internal.h"

namespace tcs {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

// ------------------------------------------------------------ profiler ----
struct Profiler {
    struct Rec { std::string name; hipEvent_t a, b; };
    std::vector<Rec> pending;
    std::vector<size_t> open;
    std::vector<hipEvent_t> pool;
    std::vector<std::string> names;              // first-seen order
    std::map<std::string, std::pair<double, long>> acc;

    hipEvent_t get() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    void begin(const char *name, hipStream_t s) {
        Rec r{name, get(), get()};
        if (!r.a || !r.b) return;
        (void)hipEventRecord(r.a, s);
        open.push_back(pending.size());
        pending.push_back(r);
    }
    void end(hipStream_t s) {
        if (open.empty()) return;
        (void)hipEventRecord(pending[open.back()].b, s);
        open.pop_back();
    }
    void flush() {
        for (auto &r : pending) {
            float ms = 0.f;
            if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                if (!acc.count(r.name)) names.push_back(r.name);
                auto &v = acc[r.name];
                v.first += ms;
                v.second += 1;
            }
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
    ~Profiler() {
        flush();
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

static thread_local Profiler *tl_prof = nullptr;

void prof_begin(const char *label, hipStream_t s) {
    if (tl_prof) tl_prof->begin(label, s);
}
void prof_end(hipStream_t s) {
    if (tl_prof) tl_prof->end(s);
}

// Makes a context's profiler current for the calling thread and times the
// enclosed stage under `label` (no-op unless profiling is enabled).
struct ProfStage {
    Profiler *prev;
    hipStream_t s;
    bool on;
    ProfStage(tcs_ctx *c, const char *label) : prev(tl_prof), s(c ? c->stream : nullptr), on(c && c->prof) {
        if (on) {
            tl_prof = c->prof;
            tl_prof->begin(label, s);
        }
    }
    ~ProfStage() {
        if (on) {
            tl_prof->end(s);
            tl_prof = prev;
        }
    }
};

// ------------------------------------------------------------ halo --------
// Single-rank mpi_subdomain_ghostcell_update (src/module_mpi_subdomain.f90:
// 758-791): the x pass copies planes i=n1m -> 0 and i=1 -> n1sub for every
// (j,k), then the z pass copies whole (i,j) planes k=n3m -> 0, k=1 -> n3sub
// (including the x ghosts just written); y is a wall (no exchange).  Both
// passes together equal: every ghost cell with i or k on a ghost plane takes
// the value at (mx(i), j, mz(k)), where mx/mz fold ghost indices onto the
// periodic interior.  Sources are never ghosts in x or z, so one kernel does
// it without ordering hazards.
template <typename E, class Set>
__global__ void k_halo(Grid g, Set fs) {
    E *__restrict__ f = fs.f[blockIdx.y];  // one field per grid row
    const int N1 = g.N1, N2 = g.N2, N3 = g.N3;
    const int n1m = g.n1m, n3m = g.n3m;
    // threads over the x-ghost faces (2*N2*N3) then the z-ghost faces (2*N1*N2)
    const size_t nx = (size_t)2 * N2 * N3, nz = (size_t)2 * N1 * N2;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nx) {
        const int side = (int)(q % 2);
        const size_t r = q / 2;
        const int j = (int)(r % N2), k = (int)(r / N2);
        if (k == 0 || k == N3 - 1) return;  // written by the z faces
        const int i = side ? n1m + 1 : 0;
        const int is = side ? 1 : n1m;
        f[g.ix(i, j, k)] = f[g.ix(is, j, k)];
    } else if (q < nx + nz) {
        const size_t r = q - nx;
        const int side = (int)(r % 2);
        const size_t t = r / 2;
        const int i = (int)(t % N1), j = (int)(t / N1);
        const int k = side ? n3m + 1 : 0;
        const int ks = side ? 1 : n3m;
        const int is = (i == 0) ? n1m : (i == n1m + 1 ? 1 : i);
        f[g.ix(i, j, k)] = f[g.ix(is, j, ks)];
    }
}

// the cells of a full field outside the interior: the faces i in {0, n1m+1}
// (all j, k), then j in {0, n2m+1} (interior i, all k), then k in {0, n3m+1}
// (interior i and j)
template <typename E, class DstSet>
__global__ void k_copy_shell(Grid g, FieldSet src_set, DstSet dst_set) {
    const double *__restrict__ src = src_set.f[blockIdx.y];  // one pair per grid row
    E *__restrict__ dst = dst_set.f[blockIdx.y];
    const int N1 = g.N1, N2 = g.N2, N3 = g.N3, n1m = g.n1m, n2m = g.n2m;
    const size_t fi = (size_t)2 * N2 * N3, fj = (size_t)2 * n1m * N3, fk = (size_t)2 * n1m * n2m;
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t o;
    if (q < fi) {
        const int side = (int)(q % 2);
        const size_t r = q / 2;
        o = g.ix(side ? N1 - 1 : 0, (int)(r % N2), (int)(r / N2));
    } else if ((q -= fi) < fj) {
        const int side = (int)(q % 2);
        const size_t r = q / 2;
        o = g.ix((int)(r % n1m) + 1, side ? N2 - 1 : 0, (int)(r / n1m));
    } else if ((q -= fj) < fk) {
        const int side = (int)(q % 2);
        const size_t r = q / 2;
        o = g.ix((int)(r % n1m) + 1, (int)(r / n1m) + 1, side ? N3 - 1 : 0);
    } else {
        return;
    }
    dst[o] = (E)src[o];
}

int copy_shell(tcs_ctx *c, std::initializer_list<std::pair<const double *, double *>> pairs) {
    FieldSet src{}, dst{};
    int nf = 0;
    for (const auto &pr : pairs) {
        if (nf == FieldSet::MAX) return TCS_E_ARG;
        src.f[nf] = const_cast<double *>(pr.first);
        dst.f[nf++] = pr.second;
    }
    const size_t n = (size_t)2 * c->N2 * c->N3 + (size_t)2 * c->n1m * c->N3 + (size_t)2 * c->n1m * c->n2m;
    TCS_LAUNCH(c->stream, "copy_shell", (k_copy_shell<double, FieldSet>), dim3((unsigned)((n + 255) / 256), (unsigned)nf),
               dim3(256), 0, c->grid(), src, dst);
    return TCS_OK;
}
int copy_shell(tcs_ctx *c, const double *src, double *dst) { return copy_shell(c, {{src, dst}}); }


int halo(tcs_ctx *c, std::initializer_list<double *> fields) {
    FieldSet fs{};
    int nf = 0;
    for (double *f : fields) {
        if (nf == FieldSet::MAX) return TCS_E_ARG;
        fs.f[nf++] = f;
    }
    const size_t n = (size_t)2 * c->N2 * c->N3 + (size_t)2 * c->N1 * c->N2;
    TCS_LAUNCH(c->stream, "halo", (k_halo<double, FieldSet>), dim3((unsigned)((n + 255) / 256), (unsigned)nf),
               dim3(256), 0, c->grid(), fs);
    return TCS_OK;
}
int halo(tcs_ctx *c, double *f) { return halo(c, {f}); }

int halo(tcs_ctx *c, std::initializer_list<float *> fields) {
    FieldSet32 fs{};
    int nf = 0;
    for (float *f : fields) {
        if (nf == FieldSet::MAX || !f) return TCS_E_ARG;
        fs.f[nf++] = f;
    }
    const size_t n = (size_t)2 * c->N2 * c->N3 + (size_t)2 * c->N1 * c->N2;
    TCS_LAUNCH(c->stream, "halo_f32", (k_halo<float, FieldSet32>), dim3((unsigned)((n + 255) / 256), (unsigned)nf),
               dim3(256), 0, c->grid(), fs);
    return TCS_OK;
}

// ------------------------------------------------- the mixed form's twins --
// the TWIN_* bit of a field's FP32 copy in the mixed form
static unsigned twin_bit(int id) {
    switch (id) {
    case TCS_T: return TWIN_T;
    case TCS_U: return TWIN_U;
    case TCS_V: return TWIN_V;
    case TCS_W: return TWIN_W;
    case TCS_P: return TWIN_P;
    case TCS_MU: case TCS_INVRHO: return TWIN_TAB;
    default: return 0;
    }
}

// The cells outside the interior of fields into their FP32 twins (the
// interior's twins written by the kernel that updated it): the twins of
// those fields are current afterwards.
static int shell_to_twins(tcs_ctx *c, std::initializer_list<int> ids) {
    FieldSet src{};
    FieldSet32 dst{};
    int nf = 0;
    unsigned bits = 0;
    for (int id : ids) {
        if (nf == FieldSet::MAX || !(dst.f[nf] = twin(c, c->f[id]))) return TCS_E_ARG;
        src.f[nf++] = c->f[id];
        bits |= twin_bit(id);
    }
    const size_t n = (size_t)2 * c->N2 * c->N3 + (size_t)2 * c->n1m * c->N3 + (size_t)2 * c->n1m * c->n2m;
    TCS_LAUNCH(c->stream, "twin_shell", (k_copy_shell<float, FieldSet32>), dim3((unsigned)((n + 255) / 256), (unsigned)nf),
               dim3(256), 0, c->grid(), src, dst);
    c->twin_ok |= bits;
    return TCS_OK;
}

__global__ void k_to_f32(const double *__restrict__ src, float *__restrict__ dst, size_t n) {
    const size_t s = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += s) dst[q] = (float)src[q];
}
__global__ void k_to_f64(const float *__restrict__ src, double *__restrict__ dst, size_t n) {
    const size_t s = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += s) dst[q] = (double)src[q];
}

static int to_f32(tcs_ctx *c, const double *src, size_t n) {
    float *dst = twin(c, src);
    if (!dst || !twin(c, src + n - 1)) {
        set_error("mixed form: an array outside the FP32-twinned arena");
        return TCS_E_ARG;
    }
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    TCS_LAUNCH(c->stream, "twin_f32", k_to_f32, dim3(nb), dim3(256), 0, src, dst, n);
    return TCS_OK;
}

// a full field's span: its front, the field, one element after
static size_t span_full(const tcs_ctx *c) { return c->nfull + FIELD_FRONT + 1; }

int ensure_twins(tcs_ctx *c, unsigned bits) {
    if (!c->a32) return TCS_OK;
    const unsigned need = bits & ~c->twin_ok;
    const struct { unsigned bit; int q; } fields[] = {
        {TWIN_T, TCS_T}, {TWIN_U, TCS_U}, {TWIN_V, TCS_V}, {TWIN_W, TCS_W}, {TWIN_P, TCS_P}};
    for (const auto &f : fields)
        if (need & f.bit) TCS_TRY(to_f32(c, c->f[f.q] - FIELD_FRONT, span_full(c)));
    if (need & TWIN_TAB) {
        TCS_TRY(to_f32(c, c->f[TCS_MU], c->ntab));
        TCS_TRY(to_f32(c, c->f[TCS_INVRHO], c->ntab));
    }
    if (need & TWIN_WALL)
        for (int w = 0; w < TCS_NWALL; w++) TCS_TRY(to_f32(c, c->wall[w], c->nwall));
    if (need & TWIN_STALE) TCS_TRY(to_f32(c, c->stale, 16));
    c->twin_ok |= bits;
    return TCS_OK;
}

static bool is_increment(int id) { return id == TCS_DU || id == TCS_DV || id == TCS_DW; }

int increment_to_fp64(tcs_ctx *c, int id) {
    if (!c->a32 || !is_increment(id)) return TCS_OK;
    const double *p = c->f[id] - FIELD_FRONT;
    const size_t n = span_full(c);
    TCS_LAUNCH(c->stream, "twin_f64", k_to_f64, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)), dim3(256), 0,
               twin(c, p), const_cast<double *>(p), n);
    return TCS_OK;
}

int increment_from_fp64(tcs_ctx *c, int id) {
    if (!c->a32 || !is_increment(id)) return TCS_OK;
    return to_f32(c, c->f[id] - FIELD_FRONT, span_full(c));
}

// ------------------------------------------------------------ reductions --
constexpr int RED_BLOCKS = 1024;

// mode 0: signed divergence (src/module_post.f90:134-150)
// mode 1: sum |u|/dx (src/module_post.f90:200-216)
template <int MODE>
__global__ void __launch_bounds__(256)
k_post_max(Grid g, const double *__restrict__ U, const double *__restrict__ V, const double *__restrict__ W,
           double *__restrict__ partial) {
    __shared__ double s[256];
    const int nrows = g.n2m * g.n3m;
    double m = -INFINITY;
    // x-rows (j,k) spread over the workgroups, i over the threads (coalesced)
    for (int r = blockIdx.x; r < nrows; r += gridDim.x) {
        const int j = r % g.n2m + 1, k = r / g.n2m + 1;
        for (int i = threadIdx.x + 1; i <= g.n1m; i += blockDim.x) {
            double v;
            if (MODE == 0)
                v = (U[g.ix(i + 1, j, k)] - U[g.ix(i, j, k)]) / g.dx1[i]
                  + (V[g.ix(i, j + 1, k)] - V[g.ix(i, j, k)]) / g.dx2[j]
                  + (W[g.ix(i, j, k + 1)] - W[g.ix(i, j, k)]) / g.dx3[k];
            else
                v = (fabs(U[g.ix(i + 1, j, k)] + U[g.ix(i, j, k)]) / 2.0) / g.dx1[i]
                  + (fabs(V[g.ix(i, j + 1, k)] + V[g.ix(i, j, k)]) / 2.0) / g.dx2[j]
                  + (fabs(W[g.ix(i, j, k + 1)] + W[g.ix(i, j, k)]) / 2.0) / g.dx3[k];
            m = fmax(m, v);
        }
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// both maxima in one pass over U, V, W: partial[b] (divergence) and
// partial[RED_BLOCKS + b] (sum |u|/dx)
__global__ void __launch_bounds__(256)
k_post_both(Grid g, const double *__restrict__ U, const double *__restrict__ V, const double *__restrict__ W,
            double *__restrict__ partial) {
    __shared__ double s0[256], s1[256];
    // Workgroups go to the 8 XCDs round-robin.  When n2m divides by 8, XCD x
    // serves the rows j of band x (n2m/8 of them) at every k, so the V(j+1)
    // row a row reads is read by a workgroup of the same XCD at about the
    // same time (one fetch into its L2, not one per XCD); otherwise the rows
    // go round-robin as they come.
    const bool banded = g.n2m % 8 == 0 && gridDim.x % 8 == 0;
    const int band = banded ? g.n2m / 8 : g.n2m, x = banded ? (int)blockIdx.x % 8 : 0;
    const int q = banded ? (int)blockIdx.x / 8 : (int)blockIdx.x, nq = banded ? (int)gridDim.x / 8 : (int)gridDim.x;
    const int nrows = band * g.n3m;
    double m0 = -INFINITY, m1 = -INFINITY;
    for (int r = q; r < nrows; r += nq) {
        const int j = x * band + r % band + 1, k = r / band + 1;
        for (int i = threadIdx.x + 1; i <= g.n1m; i += blockDim.x) {
            const double u0 = U[g.ix(i, j, k)], u1 = U[g.ix(i + 1, j, k)];
            const double v0 = V[g.ix(i, j, k)], v1 = V[g.ix(i, j + 1, k)];
            const double w0 = W[g.ix(i, j, k)], w1 = W[g.ix(i, j, k + 1)];
            m0 = fmax(m0, (u1 - u0) / g.dx1[i] + (v1 - v0) / g.dx2[j] + (w1 - w0) / g.dx3[k]);
            m1 = fmax(m1, (fabs(u1 + u0) / 2.0) / g.dx1[i] + (fabs(v1 + v0) / 2.0) / g.dx2[j]
                              + (fabs(w1 + w0) / 2.0) / g.dx3[k]);
        }
    }
    s0[threadIdx.x] = m0;
    s1[threadIdx.x] = m1;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s0[threadIdx.x] = fmax(s0[threadIdx.x], s0[threadIdx.x + h]);
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0[0];
        partial[RED_BLOCKS + blockIdx.x] = s1[0];
    }
}

__global__ void k_final_max(int n, const double *__restrict__ partial, double *__restrict__ out) {
    __shared__ double s[256];
    double m = -INFINITY;
    for (int q = threadIdx.x; q < n; q += blockDim.x) m = fmax(m, partial[q]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

template <int MODE>
static int post_max(tcs_ctx *c, double *out) {
    TCS_LAUNCH(c->stream, MODE == 0 ? "post_div" : "post_cfl", k_post_max<MODE>, dim3(RED_BLOCKS), dim3(256), 0, c->grid(), c->f[TCS_U],
                       c->f[TCS_V], c->f[TCS_W], c->red);
    TCS_LAUNCH(c->stream, "final_max", k_final_max, dim3(1), dim3(256), 0, RED_BLOCKS, c->red, c->red + RED_BLOCKS);
    TCS_HIP(hipMemcpyAsync(c->h_pinned, c->red + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TCS_HIP(hipStreamSynchronize(c->stream));
    *out = c->h_pinned[0];
    return TCS_OK;
}

// mpi_Post_Div's pass also gives mpi_Post_CFL's maximum (the same U, V, W
// reads), kept until an entry that may change the fields runs
int post_div(tcs_ctx *c, double *out) {
    TCS_LAUNCH(c->stream, "post_div_cfl", k_post_both, dim3(RED_BLOCKS), dim3(256), 0, c->grid(), c->f[TCS_U],
               c->f[TCS_V], c->f[TCS_W], c->red);
    TCS_LAUNCH(c->stream, "final_max", k_final_max, dim3(1), dim3(256), 0, RED_BLOCKS, c->red,
               c->red + 2 * RED_BLOCKS);
    TCS_LAUNCH(c->stream, "final_max", k_final_max, dim3(1), dim3(256), 0, RED_BLOCKS, c->red + RED_BLOCKS,
               c->red + 2 * RED_BLOCKS + 1);
    TCS_HIP(hipMemcpyAsync(c->h_pinned, c->red + 2 * RED_BLOCKS, 2 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    TCS_HIP(hipStreamSynchronize(c->stream));
    *out = c->h_pinned[0];
    c->cfl_max = c->h_pinned[1];
    c->cfl_valid = 1;
    return TCS_OK;
}
int post_cfl(tcs_ctx *c, double *out) {
    if (c->cfl_valid) {
        *out = c->cfl_max;
        return TCS_OK;
    }
    return post_max<1>(c, out);
}

// wall Nusselt numbers from the wall rows the solver sees (T(:,0,:) and
// T(:,n2sub,:)): pairwise sums over i per k, then over k
__global__ void k_nusselt_lines(Grid g, const double *__restrict__ T, const double *__restrict__ wbt,
                                const double *__restrict__ wup, double *__restrict__ lines) {
    extern __shared__ double buf[];
    const int k = blockIdx.x + 1;
    const int side = blockIdx.y;
    int P = 1;
    while (P < g.n1m) P <<= 1;
    for (int t = threadIdx.x; t < P; t += blockDim.x) {
        double v = 0.0;
        if (t < g.n1m) {
            const int i = t + 1;
            if (side == 0) v = (wbt[g.b2(i, k)] - T[g.ix(i, 1, k)]) / g.dmx2[1];
            else v = (T[g.ix(i, g.n2m, k)] - wup[g.b2(i, k)]) / g.dmx2[g.n2m + 1];
        }
        buf[t] = v;
    }
    __syncthreads();
    for (int h = P / 2; h >= 1; h >>= 1) {
        for (int t = threadIdx.x; t < h; t += blockDim.x) buf[t] = buf[t] + buf[t + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) lines[(size_t)side * g.n3m + (k - 1)] = buf[0];
}

__global__ void k_nusselt_total(int n1m, int n3m, double L2, double dT, const double *__restrict__ lines,
                                double *__restrict__ out) {
    extern __shared__ double buf[];
    int P = 1;
    while (P < n3m) P <<= 1;
    for (int side = 0; side < 2; side++) {
        for (int t = threadIdx.x; t < P; t += blockDim.x) buf[t] = (t < n3m) ? lines[(size_t)side * n3m + t] : 0.0;
        __syncthreads();
        for (int h = P / 2; h >= 1; h >>= 1) {
            for (int t = threadIdx.x; t < h; t += blockDim.x) buf[t] = buf[t] + buf[t + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[side] = buf[0] / (double)n1m / (double)n3m * L2 / dT;
        __syncthreads();
    }
}

int nusselt(tcs_ctx *c, double Tup, double Tbt, double *nb, double *nt) {
    int P1 = 1, P3 = 1;
    while (P1 < c->n1m) P1 <<= 1;
    while (P3 < c->n3m) P3 <<= 1;
    TCS_LAUNCH(c->stream, "nusselt_lines", k_nusselt_lines, dim3(c->n3m, 2), dim3(256), P1 * sizeof(double), c->grid(),
               c->f[TCS_T], c->wall[TCS_TBCBT], c->wall[TCS_TBCUP], c->red);
    TCS_LAUNCH(c->stream, "nusselt_total", k_nusselt_total, dim3(1), dim3(256), P3 * sizeof(double), c->n1m, c->n3m,
                       c->prm.L2, Tbt - Tup, c->red, c->red + 2 * c->n3m);
    TCS_HIP(hipMemcpyAsync(c->h_pinned, c->red + 2 * c->n3m, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TCS_HIP(hipStreamSynchronize(c->stream));
    *nb = c->h_pinned[0];
    *nt = c->h_pinned[1];
    return TCS_OK;
}

// ------------------------------------------------------ boundary rows ----
__global__ void k_fill_wall(int n, double v, double *__restrict__ w) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) w[q] = v;
}

// f(:,jrow,:) = w(:,:)
__global__ void k_set_row(Grid g, int jrow, const double *__restrict__ w, double *__restrict__ f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= g.N1) return;
    f[g.ix(i, jrow, k)] = w[g.b2(i, k)];
}

}  // namespace tcs

using namespace tcs;

// ---------------------------------------------------------------- C-ABI ---
extern "C" const char *tcs_last_error(void) { return g_err.c_str(); }

extern "C" int tcs_device_count(int *count) {
    if (!count) return TCS_E_ARG;
    TCS_HIP(hipGetDeviceCount(count));
    return TCS_OK;
}

extern "C" int tcs_mesh_axis(int32_t nm, double L, int32_t uniform, double gamma, int32_t periodic, double *x,
                             double *dx, double *dmx) {
    if (nm < 1 || !x || !dx || !dmx) {
        set_error("tcs_mesh_axis: bad arguments");
        return TCS_E_ARG;
    }
    const int n = nm + 1;
    for (int i = 0; i <= n; i++) { x[i] = 0.0; dx[i] = 0.0; dmx[i] = 0.0; }
    for (int i = 0; i <= n; i++) {
        if (uniform == 1)
            x[i] = (double)(i - 1) * L / (double)nm + 0.0;
        else
            x[i] = L * 0.5 * (1.0 + std::tanh(0.5 * gamma * (2.0 * (double)(i - 1) / (double)nm - 1.0)) / std::tanh(gamma * 0.5)) + 0.0;
    }
    if (!periodic) x[0] = x[1];
    for (int i = 1; i <= nm; i++) dx[i] = x[i + 1] - x[i];
    if (periodic) { dx[0] = dx[nm]; dx[n] = dx[1]; }
    else { dx[0] = 0.0; dx[n] = 0.0; }
    for (int i = 1; i <= nm; i++) dmx[i] = 0.5 * (dx[i - 1] + dx[i]);
    if (periodic) { dmx[0] = dmx[nm]; dmx[n] = dmx[1]; }
    else dmx[n] = 0.5 * (dx[nm] + dx[n]);
    return TCS_OK;
}

static bool pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }

static int dalloc(double **p, size_t n) {
    if (hipMalloc((void **)p, n * sizeof(double) > 0 ? n * sizeof(double) : 8) != hipSuccess) {
        set_error("hipMalloc failed (" + std::to_string(n * sizeof(double)) + " bytes)");
        *p = nullptr;
        return TCS_E_NOMEM;
    }
    if (hipMemset(*p, 0, n * sizeof(double)) != hipSuccess) {
        set_error("hipMemset failed");
        return TCS_E_HIP;
    }
    return TCS_OK;
}

// a full field: nfull elements after FIELD_FRONT, the allocation 128-byte
// aligned (hipMalloc aligns to at least 256 bytes)
static int falloc(double **p, size_t nfull) {
    double *b = nullptr;
    TCS_TRY(dalloc(&b, nfull + FIELD_FRONT + 1));
    *p = b + FIELD_FRONT;
    return TCS_OK;
}
static bool is_table(int q) { return q == TCS_INVRHO || q == TCS_MU; }
static bool in_arena(const tcs_ctx *c, const double *p) { return c->a64 && p >= c->a64 && p < c->a64 + c->an; }
static void dfree(const tcs_ctx *c, double *p) {
    if (p && !in_arena(c, p)) (void)hipFree(p);
}
static void ffree(const tcs_ctx *c, double *p) {
    if (p) dfree(c, p - FIELD_FRONT);
}

// The mixed form's arena (tcs_ctx::a64, zeroed): n elements on a 256-byte
// boundary, or a plain allocation in the FP64 forms.
static int adalloc(tcs_ctx *c, double **p, size_t n) {
    if (!c->a64) return dalloc(p, n);
    const size_t at = (c->aused + 31) / 32 * 32;
    if (at + n > c->an) {
        set_error("tcs_init: mixed-form arena too small");
        return TCS_E_NOMEM;
    }
    c->aused = at + n;
    *p = c->a64 + at;
    return TCS_OK;
}
static int afalloc(tcs_ctx *c, double **p, size_t nfull) {
    double *b = nullptr;
    TCS_TRY(adalloc(c, &b, nfull + FIELD_FRONT + 1));
    *p = b + FIELD_FRONT;
    return TCS_OK;
}
// full fields the mixed form's rows read (the state and the increments)
static bool twinned(int q) {
    return q == TCS_T || q == TCS_U || q == TCS_V || q == TCS_W || q == TCS_P || q == TCS_DU || q == TCS_DV ||
           q == TCS_DW;
}

static void free_ctx(tcs_ctx *c) {
    if (!c) return;
    adjoint_free(c);
    double **bufs[] = {&c->dx1, &c->dx2, &c->dx3, &c->dmx1, &c->dmx2, &c->dmx3, &c->rmet, &c->am, &c->ac, &c->ap, &c->rhs,
                       &c->stale, &c->prhs, &c->spec_x, &c->spec_y, &c->pA, &c->pC, &c->dxk2, &c->dzk, &c->ptab,
                       &c->tw1, &c->tw3, &c->linesum, &c->planesum, &c->scal, &c->red};
    for (double **b : bufs) dfree(c, *b);
    for (int q = 0; q < TCS_NFIELD; q++) {
        if (is_table(q)) {
            dfree(c, c->f[q]);
        } else {
            ffree(c, c->f[q]);
            ffree(c, c->spare[q]);
        }
    }
    for (int q = 0; q < TCS_NWALL; q++) dfree(c, c->wall[q]);
    if (c->a64) (void)hipFree(c->a64);
    if (c->a32) (void)hipFree(c->a32);
    if (c->jC) (void)hipFree(c->jC);
    if (c->jS) (void)hipFree(c->jS);
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
    if (c->prof) delete c->prof;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    free(c->h_dx1); free(c->h_dx2); free(c->h_dx3); free(c->h_dmx1); free(c->h_dmx2); free(c->h_dmx3);
    delete c;
}

static int upload_tables(tcs_ctx *c) {
    const int n1m = c->n1m, n2m = c->n2m, n3m = c->n3m;
    const double PI = std::acos(-1.0);
    // mpi_pressure_wave_number (src/module_solve_pressure.f90:93-129), one rank
    std::vector<double> dzk(n3m), dxk2(c->h1p);
    const double L1 = c->prm.L1, L3 = c->prm.L3;
    const double ddx = L1 / (double)n1m, ddz = L3 / (double)n3m;
    for (int k = 1; k <= n3m; k++) {
        const int km = (k <= c->h3p) ? k - 1 : n3m - k + 1;
        dzk[k - 1] = 2.0 * (1.0 - std::cos(2.0 * PI * (double)km * ddz / L3)) / (ddz * ddz);
    }
    for (int i = 1; i <= c->h1p; i++)
        dxk2[i - 1] = 2.0 * (1.0 - std::cos(2.0 * PI * (double)(i - 1) * ddx / L1)) / (ddx * ddx);
    TCS_HIP(hipMemcpy(c->dzk, dzk.data(), n3m * sizeof(double), hipMemcpyHostToDevice));
    TCS_HIP(hipMemcpy(c->dxk2, dxk2.data(), c->h1p * sizeof(double), hipMemcpyHostToDevice));
    // y coefficients of the Poisson systems (:599-611)
    std::vector<double> tab(3 * (size_t)n2m);
    for (int j = 1; j <= n2m; j++) {
        double amj = 1.0 / c->h_dx2[j] / c->h_dmx2[j];
        double apj = 1.0 / c->h_dx2[j] / c->h_dmx2[j + 1];
        if (j == 1) amj = 0.0;
        if (j == n2m) apj = 0.0;
        tab[j - 1] = amj;
        tab[n2m + j - 1] = apj;
        tab[2 * n2m + j - 1] = -amj - apj;
    }
    TCS_HIP(hipMemcpy(c->ptab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    // twiddles W_n^k = exp(-2 pi i k / n), k < n/2
    auto twid = [&](int n, double *dst) -> int {
        std::vector<double> w((size_t)n);
        for (int k = 0; k < n / 2; k++) {
            const double a = 2.0 * PI * (double)k / (double)n;
            w[2 * k] = std::cos(a);
            w[2 * k + 1] = -std::sin(a);
        }
        TCS_HIP(hipMemcpy(dst, w.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice));
        return TCS_OK;
    };
    TCS_TRY(twid(n1m, c->tw1));
    TCS_TRY(twid(n3m, c->tw3));
    return TCS_OK;
}

extern "C" int tcs_init(const tcs_params *prm, const double *dx1, const double *dx2, const double *dx3,
                        const double *dmx1, const double *dmx2, const double *dmx3, tcs_ctx **out) {
    if (!prm || !dx1 || !dx2 || !dx3 || !dmx1 || !dmx2 || !dmx3 || !out) {
        set_error("tcs_init: null argument");
        return TCS_E_ARG;
    }
    *out = nullptr;
    if (!(prm->pbc1 == 1 && prm->pbc2 == 0 && prm->pbc3 == 1)) {
        set_error("tcs_init: only periodic x/z with walls in y (pbc = T,F,T) is supported");
        return TCS_E_ARG;
    }
    if (!pow2(prm->n1m) || !pow2(prm->n3m) || prm->n1m < 4 || prm->n3m < 4 || prm->n1m > 4096 ||
        prm->n3m > 4096 || prm->n2m < 3) {
        set_error("tcs_init: n1m and n3m must be powers of two in [4, 4096] and n2m >= 3");
        return TCS_E_ARG;
    }
    if (prm->precision < 0 || prm->precision > 2) {
        set_error("tcs_init: precision must be 0 (plain FP64, reference operation order), 1 (plain FP64, "
                  "contracted/reciprocal arithmetic) or 2 (mixed: FP32 ADI scratch and line solves)");
        return TCS_E_ARG;
    }
    if (prm->precision == 2 && (prm->n1m > 512 || prm->n2m > 512 || prm->n3m > 512)) {
        set_error("tcs_init: the mixed form solves its lines split in parts and needs n1m, n2m, n3m <= 512");
        return TCS_E_ARG;
    }
    if (prm->tdma_split != 0 && prm->tdma_split != 1) {
        set_error("tcs_init: tdma_split must be 0 (single-rank line solves) or 1 (split lines)");
        return TCS_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("tcs_init: no HIP device");
        return TCS_E_HIP;
    }
    tcs_ctx *c = new tcs_ctx();
    std::memset((void *)c, 0, sizeof(tcs_ctx));
    c->prm = *prm;
    Phys &p = c->phys;
    p.Ct = prm->Ct; p.Cmu = prm->Cmu; p.Cmt = prm->Cmt; p.Cmp = prm->Cmp; p.DeltaT = prm->DeltaT;
    p.a10 = prm->a10; p.a11 = prm->a11; p.a12 = prm->a12; p.a13 = prm->a13; p.a14 = prm->a14; p.a15 = prm->a15;
    p.b10 = prm->b10; p.b11 = prm->b11; p.b12 = prm->b12; p.b13 = prm->b13; p.b14 = prm->b14; p.b15 = prm->b15;
    p.c10 = prm->c10; p.c11 = prm->c11; p.c12 = prm->c12; p.c13 = prm->c13; p.c14 = prm->c14; p.c15 = prm->c15;
    p.d10 = prm->d10; p.d11 = prm->d11; p.d12 = prm->d12; p.d13 = prm->d13; p.d14 = prm->d14; p.d15 = prm->d15;
    p.KPP0 = prm->KPP0; p.Rho0 = prm->Rho0; p.Cp0 = prm->Cp0; p.Mu0 = prm->Mu0; p.a12pera11 = prm->a12pera11;
    p.kppS = p.c10 / p.KPP0;
    p.rcS = p.Rho0 * p.Cp0;
    // the reference's expression of the d(1/(rho cp)) numerator
    // (module_solve_thermal.f90:116-156) collected by powers of DeltaT*T
    p.dircN[0] = p.a11 + p.b11;
    p.dircN[1] = 2.0 * p.a11 * p.b11 + 2.0 * p.a12 + 2.0 * p.b12;
    p.dircN[2] = 3.0 * p.a12 * p.b11 + 3.0 * p.a11 * p.b12 + 3.0 * p.a13 + 3.0 * p.b13;
    p.dircN[3] = 4.0 * p.a13 * p.b11 + 4.0 * p.a11 * p.b13 + 4.0 * p.a12 * p.b12;
    p.dircN[4] = 5.0 * p.a13 * p.b12 + 5.0 * p.a12 * p.b13;
    p.dircN[5] = 6.0 * p.a13 * p.b13;

    c->n1m = prm->n1m; c->n2m = prm->n2m; c->n3m = prm->n3m;
    c->tdma_split = prm->tdma_split;
    c->N1 = c->n1m + 2; c->N2 = c->n2m + 2; c->N3 = c->n3m + 2;
    c->P1 = field_pitch(c->N1);
    c->nfull = (size_t)c->P1 * c->N2 * c->N3;
    c->nint = (size_t)c->n1m * c->n2m * c->n3m;
    c->nwall = (size_t)c->N1 * c->N3;
    c->ntab = (size_t)c->N2 * c->N3;
    c->h1p = c->n1m / 2 + 1;
    c->h3p = c->n3m / 2 + 1;
    int rc = TCS_OK;
    auto fail = [&](int r) { free_ctx(c); return r; };
    if (hipGetDevice(&c->device) != hipSuccess) { set_error("hipGetDevice failed"); return fail(TCS_E_HIP); }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(TCS_E_HIP);
    }
    auto hcopy = [](const double *src, int n) {
        double *d = (double *)malloc(sizeof(double) * n);
        std::memcpy(d, src, sizeof(double) * n);
        return d;
    };
    c->h_dx1 = hcopy(dx1, c->N1); c->h_dmx1 = hcopy(dmx1, c->N1);
    c->h_dx2 = hcopy(dx2, c->N2); c->h_dmx2 = hcopy(dmx2, c->N2);
    c->h_dx3 = hcopy(dx3, c->N3); c->h_dmx3 = hcopy(dmx3, c->N3);

    struct { double **p; size_t n; } allocs[] = {
        {&c->dx1, (size_t)c->N1}, {&c->dmx1, (size_t)c->N1}, {&c->dx2, (size_t)c->N2}, {&c->dmx2, (size_t)c->N2},
        {&c->dx3, (size_t)c->N3}, {&c->dmx3, (size_t)c->N3},
        {&c->rmet, (size_t)4 * (c->N1 + c->N2 + c->N3)},
        {&c->am, c->nint}, {&c->ac, c->nint}, {&c->ap, c->nint}, {&c->rhs, c->nint},
        {&c->prhs, c->nint},
        {&c->spec_x, (size_t)2 * c->h1p * c->n2m * c->n3m},
        {&c->pA, (size_t)c->h1p * c->n2m * c->n3m}, {&c->pC, (size_t)c->h1p * c->n2m * c->n3m},
        {&c->dxk2, (size_t)c->h1p}, {&c->dzk, (size_t)c->n3m}, {&c->ptab, (size_t)3 * c->n2m},
        {&c->tw1, (size_t)c->n1m}, {&c->tw3, (size_t)c->n3m},
        {&c->linesum, (size_t)c->n2m * c->n3m}, {&c->planesum, (size_t)c->n3m}, {&c->scal, 8},
        {&c->red, (size_t)2 * RED_BLOCKS + 16 + (size_t)4 * c->n3m},
    };
    for (auto &a : allocs)
        if ((rc = dalloc(a.p, a.n)) != TCS_OK) return fail(rc);
    if (prm->precision == 2) {
        // the arrays the mixed form's rows and block corrections read, in one
        // arena with an FP32 twin of each (tcs_ctx::a64, a32)
        auto up = [](size_t n) { return (n + 31) / 32 * 32; };
        c->an = 12 * up(c->nfull + FIELD_FRONT + 1) + 2 * up(c->ntab) + TCS_NWALL * up(c->nwall) + up(16);
        if (hipMalloc((void **)&c->a64, c->an * sizeof(double)) != hipSuccess ||
            hipMalloc((void **)&c->a32, c->an * sizeof(float)) != hipSuccess) {
            set_error("hipMalloc failed (mixed-form arenas, " + std::to_string(c->an * 12) + " bytes)");
            return fail(TCS_E_NOMEM);
        }
        if (hipMemset(c->a64, 0, c->an * sizeof(double)) != hipSuccess ||
            hipMemset(c->a32, 0, c->an * sizeof(float)) != hipSuccess) {
            set_error("hipMemset failed");
            return fail(TCS_E_HIP);
        }
    }
    if ((rc = adalloc(c, &c->stale, 16)) != TCS_OK) return fail(rc);
    for (int q = 0; q < TCS_NFIELD; q++) {
        if (is_table(q)) rc = adalloc(c, &c->f[q], c->ntab);
        else rc = twinned(q) ? afalloc(c, &c->f[q], c->nfull) : falloc(&c->f[q], c->nfull);
        if (rc != TCS_OK) return fail(rc);
    }
    for (int q : {TCS_T, TCS_U, TCS_V, TCS_W, TCS_DPHAT, TCS_DP0}) {
        rc = twinned(q) ? afalloc(c, &c->spare[q], c->nfull) : falloc(&c->spare[q], c->nfull);
        if (rc != TCS_OK) return fail(rc);
    }
    for (int q = 0; q < TCS_NWALL; q++)
        if ((rc = adalloc(c, &c->wall[q], c->nwall)) != TCS_OK) return fail(rc);
    if (hipMalloc((void **)&c->jC, c->N2 * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&c->jS, c->N2 * sizeof(int)) != hipSuccess) {
        set_error("hipMalloc failed");
        return fail(TCS_E_NOMEM);
    }
    if (hipHostMalloc((void **)&c->h_pinned, 64, hipHostMallocDefault) != hipSuccess) {
        set_error("hipHostMalloc failed");
        return fail(TCS_E_NOMEM);
    }
    // masks (src/module_mpi_subdomain.f90:304-374), pbc2 = F, one rank
    std::vector<int> jC(c->N2, 1), jS(c->N2, 1);
    jC[0] = 0; jS[0] = 0; jS[1] = 0;
    jC[c->n2m + 1] = 0; jS[c->n2m + 1] = 0;
    if (hipMemcpy(c->jC, jC.data(), c->N2 * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->jS, jS.data(), c->N2 * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->dx1, dx1, c->N1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->dmx1, dmx1, c->N1 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->dx2, dx2, c->N2 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->dmx2, dmx2, c->N2 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->dx3, dx3, c->N3 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->dmx3, dmx3, c->N3 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("tcs_init: metric upload failed");
        return fail(TCS_E_HIP);
    }
    {
        // reciprocal metric tables of the fast forms (x / dx as x * (1/dx)),
        // then the tables and their reciprocals rounded to float
        std::vector<double> r;
        const std::pair<const double *, int> tabs[] = {{dx1, c->N1}, {dmx1, c->N1}, {dx2, c->N2},
                                                       {dmx2, c->N2}, {dx3, c->N3}, {dmx3, c->N3}};
        for (const auto &m : tabs)
            for (int q = 0; q < m.second; q++) r.push_back(1.0 / m.first[q]);
        const size_t nm = r.size();
        std::vector<float> f;
        for (const auto &m : tabs)
            for (int q = 0; q < m.second; q++) f.push_back((float)m.first[q]);
        for (size_t q = 0; q < nm; q++) f.push_back((float)r[q]);
        r.resize(2 * nm);
        std::memcpy(r.data() + nm, f.data(), f.size() * sizeof(float));
        if (hipMemcpy(c->rmet, r.data(), r.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("tcs_init: metric upload failed");
            return fail(TCS_E_HIP);
        }
    }
    if ((rc = upload_tables(c)) != TCS_OK) return fail(rc);
    if (hipDeviceSynchronize() != hipSuccess) {
        set_error("tcs_init: device error");
        return fail(TCS_E_HIP);
    }
    *out = c;
    return TCS_OK;
}

extern "C" int tcs_finalize(tcs_ctx *c) {
    if (!c) return TCS_E_ARG;
    c->pend = 0;  // held launches of a context being destroyed are dropped
    (void)hipStreamSynchronize(c->stream);
    free_ctx(c);
    return TCS_OK;
}

extern "C" int tcs_sync(tcs_ctx *c) {
    if (!c) return TCS_E_ARG;
    TCS_TRY(flush_pending(c));
    TCS_HIP(hipStreamSynchronize(c->stream));
    return TCS_OK;
}

// the context's stream: what callers enqueue on it follows the held launches
extern "C" void *tcs_stream(tcs_ctx *c) {
    if (!c || flush_pending(c) != TCS_OK) return nullptr;
    return (void *)c->stream;
}

static int check_field(tcs_ctx *c, int id) {
    if (!c || id < 0 || id >= TCS_NFIELD) {
        set_error("bad context or field id");
        return TCS_E_ARG;
    }
    TCS_TRY(flush_pending(c));
    c->cfl_valid = 0;  // the field may change (upload, halo, a handed-out pointer)
    c->twin_ok &= ~twin_bit(id);  // and with it the mixed form's FP32 copy
    return TCS_OK;
}

extern "C" int tcs_field_upload(tcs_ctx *c, int32_t id, const double *host) {
    TCS_TRY(check_field(c, id));
    if (!host) return TCS_E_ARG;
    if (id == TCS_INVRHO || id == TCS_MU) {
        // the reference's arrays are row-broadcast: keep the i = n1sub column
        std::vector<double> t(c->ntab);
        for (size_t r = 0; r < c->ntab; r++) t[r] = host[(size_t)(c->N1 - 1) + (size_t)c->N1 * r];
        TCS_HIP(hipMemcpyAsync(c->f[id], t.data(), c->ntab * sizeof(double), hipMemcpyHostToDevice, c->stream));
        TCS_HIP(hipStreamSynchronize(c->stream));
        return TCS_OK;
    }
    // one contiguous transfer into a second buffer (free outside tcs_step),
    // then the rows to their pitch on the device
    const size_t row = (size_t)c->N1 * sizeof(double), pitch = (size_t)c->P1 * sizeof(double);
    double *stage = c->spare[TCS_T];
    TCS_HIP(hipMemcpyAsync(stage, host, row * c->N2 * c->N3, hipMemcpyHostToDevice, c->stream));
    TCS_HIP(hipMemcpy2DAsync(c->f[id], pitch, stage, row, row, (size_t)c->N2 * c->N3, hipMemcpyDeviceToDevice,
                             c->stream));
    TCS_TRY(increment_from_fp64(c, id));  // the mixed form keeps dU, dV, dW in FP32
    TCS_HIP(hipStreamSynchronize(c->stream));
    return TCS_OK;
}

extern "C" int tcs_field_download(tcs_ctx *c, int32_t id, double *host) {
    TCS_TRY(check_field(c, id));
    if (!host) return TCS_E_ARG;
    if (id == TCS_INVRHO || id == TCS_MU) {
        std::vector<double> t(c->ntab);
        TCS_HIP(hipMemcpyAsync(t.data(), c->f[id], c->ntab * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        TCS_HIP(hipStreamSynchronize(c->stream));
        for (size_t r = 0; r < c->ntab; r++)
            for (int i = 0; i < c->N1; i++) host[(size_t)i + (size_t)c->N1 * r] = t[r];
        return TCS_OK;
    }
    const size_t row = (size_t)c->N1 * sizeof(double), pitch = (size_t)c->P1 * sizeof(double);
    double *stage = c->spare[TCS_T];
    TCS_TRY(increment_to_fp64(c, id));  // the mixed form keeps dU, dV, dW in FP32
    TCS_HIP(hipMemcpy2DAsync(stage, row, c->f[id], pitch, row, (size_t)c->N2 * c->N3, hipMemcpyDeviceToDevice,
                             c->stream));
    TCS_HIP(hipMemcpyAsync(host, stage, row * c->N2 * c->N3, hipMemcpyDeviceToHost, c->stream));
    TCS_HIP(hipStreamSynchronize(c->stream));
    return TCS_OK;
}

extern "C" int tcs_field_device_ptr(tcs_ctx *c, int32_t id, double **dptr) {
    TCS_TRY(check_field(c, id));
    if (!dptr) return TCS_E_ARG;
    TCS_TRY(increment_to_fp64(c, id));  // the mixed form keeps dU, dV, dW in FP32
    *dptr = c->f[id];
    return TCS_OK;
}

extern "C" int tcs_field_copy_device(tcs_ctx *c, int32_t id, int32_t to_field, double *dev) {
    TCS_TRY(check_field(c, id));
    if (!dev) return TCS_E_ARG;
    const size_t n = (id == TCS_INVRHO || id == TCS_MU) ? c->ntab : c->nfull;
    if (to_field) {
        TCS_HIP(hipMemcpyAsync(c->f[id], dev, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        TCS_TRY(increment_from_fp64(c, id));  // the mixed form keeps dU, dV, dW in FP32
    } else {
        TCS_TRY(increment_to_fp64(c, id));
        TCS_HIP(hipMemcpyAsync(dev, c->f[id], n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    return TCS_OK;
}

extern "C" int tcs_fields_touched(tcs_ctx *c) {
    if (!c) return TCS_E_ARG;
    TCS_TRY(flush_pending(c));
    c->cfl_valid = 0;
    c->twin_ok = 0;
    for (int q : {TCS_DU, TCS_DV, TCS_DW}) TCS_TRY(increment_from_fp64(c, q));
    return TCS_OK;
}

extern "C" int tcs_field_pitch(tcs_ctx *c, int32_t *pitch) {
    if (!c || !pitch) return TCS_E_ARG;
    *pitch = c->P1;
    return TCS_OK;
}

extern "C" int tcs_wall_upload(tcs_ctx *c, int32_t id, const double *host) {
    if (!c || id < 0 || id >= TCS_NWALL || !host) return TCS_E_ARG;
    TCS_TRY(flush_pending(c));
    c->twin_ok &= ~(unsigned)TWIN_WALL;
    TCS_HIP(hipMemcpyAsync(c->wall[id], host, c->nwall * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TCS_HIP(hipStreamSynchronize(c->stream));
    return TCS_OK;
}

extern "C" int tcs_wall_download(tcs_ctx *c, int32_t id, double *host) {
    if (!c || id < 0 || id >= TCS_NWALL || !host) return TCS_E_ARG;
    TCS_TRY(flush_pending(c));
    TCS_HIP(hipMemcpyAsync(host, c->wall[id], c->nwall * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TCS_HIP(hipStreamSynchronize(c->stream));
    return TCS_OK;
}

static int fill_wall(tcs_ctx *c, int id, double v) {
    TCS_LAUNCH(c->stream, "fill_wall", k_fill_wall, dim3((unsigned)((c->nwall + 255) / 256)), dim3(256), 0,
                       (int)c->nwall, v, c->wall[id]);
    return TCS_OK;
}

static int set_row(tcs_ctx *c, int field, int jrow, int wall) {
    TCS_LAUNCH(c->stream, "set_row", k_set_row, dim3((c->N1 + 63) / 64, c->N3), dim3(64), 0, c->grid(), jrow,
                       c->wall[wall], c->f[field]);
    return TCS_OK;
}

extern "C" int tcs_thermal_boundary(tcs_ctx *c, double Tup, double Tbt) {
    if (!c) return TCS_E_ARG;
    TCS_TRY(flush_pending(c));
    c->twin_ok = 0;
    TCS_TRY(fill_wall(c, TCS_TBCUP, Tup));
    TCS_TRY(fill_wall(c, TCS_TBCBT, Tbt));
    // comm_1d_x2%myrank == 0 branch only (single y-rank)
    TCS_TRY(set_row(c, TCS_T, 0, TCS_TBCBT));
    return TCS_OK;
}

extern "C" int tcs_momentum_boundary(tcs_ctx *c, double Uup, double Ubt, double Vup, double Vbt, double Wup,
                                     double Wbt) {
    if (!c) return TCS_E_ARG;
    TCS_TRY(flush_pending(c));
    c->cfl_valid = 0;
    c->twin_ok = 0;
    TCS_TRY(fill_wall(c, TCS_UBCUP, Uup));
    TCS_TRY(fill_wall(c, TCS_VBCUP, Vup));
    TCS_TRY(fill_wall(c, TCS_WBCUP, Wup));
    TCS_TRY(fill_wall(c, TCS_UBCBT, Ubt));
    TCS_TRY(fill_wall(c, TCS_VBCBT, Vbt));
    TCS_TRY(fill_wall(c, TCS_WBCBT, Wbt));
    TCS_TRY(set_row(c, TCS_U, 0, TCS_UBCBT));
    TCS_TRY(set_row(c, TCS_W, 0, TCS_WBCBT));
    TCS_TRY(set_row(c, TCS_V, 0, TCS_VBCBT));
    TCS_TRY(set_row(c, TCS_V, 1, TCS_VBCBT));
    return TCS_OK;
}

#define CTX_CHECK(c) \
    if (!(c)) { set_error("null context"); return TCS_E_ARG; } \
    TCS_TRY(flush_pending(c));
// entries that may change fields drop the kept CFL maximum
#define CTX_GUARD(c) \
    CTX_CHECK(c) \
    (c)->cfl_valid = 0;

#define STAGE(label) ProfStage _stage(c, label)

// the stage routine of the context's arithmetic form (precision 0 / 1 / 2)
#define FORM(c, fn, ...)                                                     \
    ((c)->prm.precision == 2   ? ::tcs::mixed::fn((c), ##__VA_ARGS__)        \
     : (c)->prm.precision == 1 ? ::tcs::fast::fn((c), ##__VA_ARGS__)         \
                               : ::tcs::exact::fn((c), ##__VA_ARGS__))

namespace tcs {
// the ghost-cell update of a field as the context's form keeps it: the FP32
// twin for the mixed form's increments dU, dV, dW
static int halo_field(tcs_ctx *c, int id) {
    if (c->a32 && is_increment(id)) return halo(c, {twin(c, c->f[id])});
    return halo(c, c->f[id]);
}

int flush_pending(tcs_ctx *c) {
    const int p = c->pend;
    if (!p) return TCS_OK;
    c->pend = 0;
    if (p & PEND_LDV) {
        ProfStage _s(c, "momentum blockLdV");
        TCS_TRY(FORM(c, momentum_blockLdV, c->pend_dt_ldv));
    }
    if (p & PEND_HALO_DV) {
        ProfStage _s(c, "update ghostcell");
        TCS_TRY(halo_field(c, TCS_DV));
    }
    if (p & PEND_LDU) {
        ProfStage _s(c, "momentum blockLdU");
        TCS_TRY(FORM(c, momentum_blockLdU, c->pend_dt_ldu));
    }
    if (p & PEND_RHS) {
        ProfStage _s(c, "pressure RHS");
        TCS_TRY(FORM(c, pressure_rhs, c->pend_dt_rhs));
    }
    return TCS_OK;
}
}  // namespace tcs

// stage labels follow the reference's timer labels (src/main.f90:55-72)
extern "C" int tcs_thermal_coeffi(tcs_ctx *c) { CTX_GUARD(c); STAGE("thermal coeff"); return FORM(c, thermal_coeffi); }
// the last sweep writes T + dT into T's second buffer, which becomes T
extern "C" int tcs_thermal_solver(tcs_ctx *c, double dt) {
    CTX_GUARD(c);
    STAGE("thermal solve");
    bool swap_T = false;
    TCS_TRY(FORM(c, thermal_solver_into, dt, c->spare[TCS_T], &swap_T));
    if (swap_T) std::swap(c->f[TCS_T], c->spare[TCS_T]);
    return TCS_OK;
}
extern "C" int tcs_momentum_coeffi(tcs_ctx *c) { CTX_GUARD(c); STAGE("momentum coeff"); return FORM(c, momentum_coeffi); }
extern "C" int tcs_momentum_solvedU(tcs_ctx *c, double dt) { CTX_GUARD(c); STAGE("momentum solvedU"); return FORM(c, momentum_solvedU, dt); }
extern "C" int tcs_momentum_solvedV(tcs_ctx *c, double dt) { CTX_GUARD(c); STAGE("momentum solvedV"); return FORM(c, momentum_solvedV, dt); }
extern "C" int tcs_momentum_solvedW(tcs_ctx *c, double dt) { CTX_GUARD(c); STAGE("momentum solvedW"); return FORM(c, momentum_solvedW, dt); }
// blockLdV waits: main follows it with the dV halo, blockLdU and the
// pseudo-update, which then run as tcs_step runs them
extern "C" int tcs_momentum_blockLdV(tcs_ctx *c, double dt) {
    CTX_GUARD(c);
    c->pend = PEND_LDV;
    c->pend_dt_ldv = dt;
    return TCS_OK;
}
extern "C" int tcs_momentum_blockLdU(tcs_ctx *c, double dt) {
    if (c && c->pend == (PEND_LDV | PEND_HALO_DV) && dt == c->pend_dt_ldv) {
        c->cfl_valid = 0;
        c->pend |= PEND_LDU;
        c->pend_dt_ldu = dt;
        return TCS_OK;
    }
    CTX_GUARD(c);
    STAGE("momentum blockLdU");
    return FORM(c, momentum_blockLdU, dt);
}
extern "C" int tcs_momentum_pseudoupdateUVW(tcs_ctx *c) {
    if (c && c->pend == (PEND_LDV | PEND_HALO_DV | PEND_LDU)) {
        // the block corrections write U + dU, V + dV, W + dW into the
        // second buffers (they read the old U, V) as in tcs_step
        c->pend = 0;
        c->cfl_valid = 0;
        {
            STAGE("momentum blockLdV");
            TCS_TRY(FORM(c, momentum_blockLdV_pseudo, c->pend_dt_ldv, c->spare[TCS_V]));
        }
        {
            STAGE("update ghostcell");
            TCS_TRY(halo_field(c, TCS_DV));
        }
        {
            STAGE("momentum blockLdU");
            TCS_TRY(FORM(c, momentum_blockLdU_pseudo, c->pend_dt_ldu, c->spare[TCS_U], c->spare[TCS_W]));
        }
        for (int q : {TCS_U, TCS_V, TCS_W}) std::swap(c->f[q], c->spare[q]);
        return TCS_OK;
    }
    CTX_GUARD(c);
    STAGE("momentum updatePseudoUVW");
    return FORM(c, momentum_pseudoupdate);
}
extern "C" int tcs_pressure_wave_number(tcs_ctx *c) { CTX_GUARD(c); return upload_tables(c); }
// the right-hand side waits for the Poisson solve, whose x transform makes it
extern "C" int tcs_pressure_rhs(tcs_ctx *c, double dt) {
    CTX_GUARD(c);
    c->pend = PEND_RHS;
    c->pend_dt_rhs = dt;
    return TCS_OK;
}
extern "C" int tcs_pressure_poisson_fft2(tcs_ctx *c) {
    if (c && c->pend == PEND_RHS) {
        c->pend = 0;
        c->cfl_valid = 0;
        STAGE("pressure PoissonSolve");
        return FORM(c, pressure_poisson_rhs, c->pend_dt_rhs);
    }
    CTX_GUARD(c);
    STAGE("pressure PoissonSolve");
    return FORM(c, pressure_poisson);
}
// the dPhat / dP0 history goes into the second buffers, which become them
extern "C" int tcs_pressure_projection(tcs_ctx *c, double dt) {
    CTX_GUARD(c);
    STAGE("pressure projection");
    TCS_TRY(FORM(c, pressure_projection_into, dt, c->spare[TCS_DPHAT], c->spare[TCS_DP0]));
    std::swap(c->f[TCS_DPHAT], c->spare[TCS_DPHAT]);
    std::swap(c->f[TCS_DP0], c->spare[TCS_DP0]);
    return TCS_OK;
}

extern "C" int tcs_ghostcell_update(tcs_ctx *c, int32_t id) {
    if (c && id == TCS_DV && c->pend == PEND_LDV) {  // the dV halo between blockLdV and blockLdU
        c->cfl_valid = 0;
        c->pend |= PEND_HALO_DV;
        return TCS_OK;
    }
    TCS_TRY(check_field(c, id));
    if (id == TCS_INVRHO || id == TCS_MU) return TCS_OK;  // tables carry no x/z ghosts of their own
    STAGE("update ghostcell");
    return halo_field(c, id);
}

extern "C" int tcs_post_div(tcs_ctx *c, double *maxdiv) {
    CTX_CHECK(c);
    if (!maxdiv) return TCS_E_ARG;
    STAGE("post div");
    return post_div(c, maxdiv);
}

extern "C" int tcs_post_cfl(tcs_ctx *c, double *newcfl, double *newdt) {
    CTX_CHECK(c);
    if (!newcfl || !newdt) return TCS_E_ARG;
    double mx = 0.0;
    {
        STAGE("post CFL");
        TCS_TRY(post_cfl(c, &mx));
    }
    // src/module_post.f90:223-226
    double nd = c->prm.MaxCFL / mx;
    if (nd > c->prm.dtMax) nd = c->prm.dtMax;
    *newdt = nd;
    *newcfl = nd * mx;
    return TCS_OK;
}

extern "C" int tcs_nusselt(tcs_ctx *c, double Tup, double Tbt, double *nb, double *nt) {
    CTX_CHECK(c);
    if (!nb || !nt) return TCS_E_ARG;
    return nusselt(c, Tup, Tbt, nb, nt);
}

// src/main.f90:131-203 (thermal -> momentum -> pressure with main's halos)
#define RUN(label, expr)              \
    do {                              \
        ProfStage _s(c, label);       \
        TCS_TRY(expr);                \
    } while (0)

extern "C" int tcs_step(tcs_ctx *c, double dt) {
    CTX_GUARD(c);
    STAGE("step");
    const char *H = "update ghostcell";
    bool swap_T = false;
    RUN("thermal solve", FORM(c, thermal_solver_into, dt, c->spare[TCS_T], &swap_T));
    if (swap_T) std::swap(c->f[TCS_T], c->spare[TCS_T]);
    RUN(H, halo(c, c->f[TCS_T]));
    // the mixed form: the last thermal sweep wrote T's FP32 copy inside; now its shell
    if (c->a32 && swap_T) RUN(H, shell_to_twins(c, {TCS_T}));
    RUN("momentum coeff", FORM(c, momentum_coeffi));
    RUN("momentum solvedU", FORM(c, momentum_solvedU, dt));
    RUN(H, halo_field(c, TCS_DU));
    RUN("momentum solvedV", FORM(c, momentum_solvedV, dt));
    RUN(H, halo_field(c, TCS_DV));
    RUN("momentum solvedW", FORM(c, momentum_solvedW, dt));
    RUN(H, halo_field(c, TCS_DW));
    // block corrections with the pseudo-update (U + dU, V + dV, W + dW into
    // the second buffers; the corrections read the old U, V)
    RUN("momentum blockLdV", FORM(c, momentum_blockLdV_pseudo, dt, c->spare[TCS_V]));
    RUN(H, halo_field(c, TCS_DV));
    RUN("momentum blockLdU", FORM(c, momentum_blockLdU_pseudo, dt, c->spare[TCS_U], c->spare[TCS_W]));
    for (int q : {TCS_U, TCS_V, TCS_W}) std::swap(c->f[q], c->spare[q]);
    RUN(H, halo(c, {c->f[TCS_U], c->f[TCS_V], c->f[TCS_W]}));
    // the right-hand side is made inside the solve's x transform
    RUN("pressure PoissonSolve", FORM(c, pressure_poisson_rhs, dt));
    RUN(H, halo(c, c->f[TCS_DP]));
    RUN("pressure projection", FORM(c, pressure_projection_into, dt, c->spare[TCS_DPHAT], c->spare[TCS_DP0]));
    std::swap(c->f[TCS_DPHAT], c->spare[TCS_DPHAT]);
    std::swap(c->f[TCS_DP0], c->spare[TCS_DP0]);
    RUN(H, halo(c, {c->f[TCS_U], c->f[TCS_V], c->f[TCS_W], c->f[TCS_P], c->f[TCS_DPHAT]}));
    // the mixed form: the projection wrote the FP32 copies of U, V, W, P inside
    if (c->a32) RUN(H, shell_to_twins(c, {TCS_U, TCS_V, TCS_W, TCS_P}));
    return TCS_OK;
}

// ---- profiling -------------------------------------------------------------
extern "C" int tcs_profile(tcs_ctx *c, int32_t enable) {
    CTX_GUARD(c);
    if (c->prof) {
        delete c->prof;
        c->prof = nullptr;
    }
    if (enable) c->prof = new Profiler();
    return TCS_OK;
}

extern "C" int tcs_profile_count(tcs_ctx *c, int32_t *n) {
    CTX_GUARD(c);
    if (!n) return TCS_E_ARG;
    if (!c->prof) { *n = 0; return TCS_OK; }
    c->prof->flush();
    *n = (int32_t)c->prof->names.size();
    return TCS_OK;
}

extern "C" int tcs_profile_entry(tcs_ctx *c, int32_t idx, char *name, int32_t name_len, double *total_ms,
                                 int64_t *count) {
    CTX_GUARD(c);
    if (!c->prof || idx < 0 || idx >= (int32_t)c->prof->names.size() || !name || name_len < 1 || !total_ms || !count)
        return TCS_E_ARG;
    const std::string &nm = c->prof->names[idx];
    std::snprintf(name, (size_t)name_len, "%s", nm.c_str());
    const auto &v = c->prof->acc[nm];
    *total_ms = v.first;
    *count = v.second;
    return TCS_OK;
}
