// psst_device.hip -- the C-ABI of include/psst_gpu.h: device contexts,
// runs, batches of instances, the precision switch and the autotuner.  The
// kernels and their launch templates are in psst_device.hpp; the launch
// templates for each (dimension, state type) pair are instantiated in
// psst_inst.hip and only declared here.
#include "psst_device.hpp"

#include <cstdlib>
#include <mutex>

#include "ordered_sched.hpp"

using namespace psst_dev;

PSST_DIM_T_ENTRY_POINTS(extern, 2, double)
PSST_DIM_T_ENTRY_POINTS(extern, 2, float)
PSST_DIM_T_ENTRY_POINTS(extern, 3, double)
PSST_DIM_T_ENTRY_POINTS(extern, 3, float)

namespace {

thread_local std::string g_create_error;

// Typical cell spacing per axis: the mean |dx_k| over interior faces whose
// centre-to-centre vector is dominated by axis k (scales the Morton curve).
void axis_spacing(const psst_mesh_desc *m, double axis_h[3]) {
    const int dim = m->dim, N = m->n_cells, F = m->n_faces;
    const int64_t NT = (int64_t)m->n_cells + m->n_ghosts;
    const int32_t *f0 = m->face_cells, *f1 = m->face_cells + F;
    double sum[3] = {0, 0, 0};
    int64_t cnt[3] = {0, 0, 0};
    for (int f = 0; f < F; ++f) {
        if (f0[f] >= N || f1[f] >= N) continue;
        int kk = 0;
        double best = -1.0;
        for (int k = 0; k < dim; ++k) {
            double d = m->cell_centre[(size_t)k * NT + f0[f]] - m->cell_centre[(size_t)k * NT + f1[f]];
            if (m->face_shift) d -= m->face_shift[(size_t)k * F + f];
            if (std::abs(d) > best) {
                best = std::abs(d);
                kk = k;
            }
        }
        sum[kk] += best;
        cnt[kk]++;
    }
    for (int k = 0; k < 3; ++k) axis_h[k] = (k < dim && cnt[k]) ? sum[k] / (double)cnt[k] : 0.0;
}

// Cells per chunk of the chunked ordered sweep: PSST_SWEEP_CHUNK_CELLS if
// set (a tuning knob), else 3,600 for one-wave workgroups and 8,192 for
// larger ones; at least N / capacity, so that the chunks fit one cooperative
// launch (capacity 0: no limit known).  On the 4.9M-cell wedge, one-wave
// workgroups: boxes of 3,000, 3,200, 3,600 and 4,096 cells swept a stage in
// 12.3, 12.6, 12.3-12.9 and 12.6 ms; 128 threads and 8,192 cells, 14.3 ms
// (earlier builds with 256 threads: 481 boxes 15.2 ms, 300 16.5 ms, 150
// 19.1 ms).
int32_t sweep_chunk_cells(int32_t N, int capacity, int threads) {
    int32_t k = threads == 64 ? 3600 : 8192;
    if (const char *e = std::getenv("PSST_SWEEP_CHUNK_CELLS")) {
        const long v = std::strtol(e, nullptr, 10);
        if (v > 0 && v <= (1L << 30)) k = (int32_t)v;
    }
    return capacity > 0 ? std::max<int32_t>(k, (int32_t)(((int64_t)N + capacity - 1) / capacity)) : k;
}

// Workgroup size of the chunked sweep: PSST_SWEEP_THREADS (64, 128 or 256;
// a tuning knob) if set, else 0: the first of 64, 128 and 256 whose chunks
// keep every entry within one pass of the workgroup.
int sweep_threads() {
    if (const char *e = std::getenv("PSST_SWEEP_THREADS")) {
        const long v = std::strtol(e, nullptr, 10);
        if (v == 64 || v == 128 || v == 256) return (int)v;
    }
    return 0;
}

// Levels between publications of a chunk's progress:
// PSST_SWEEP_PUBLISH_STRIDE if set (a tuning knob), else 1 (after every
// waited-for level).  On the 4.9M-cell wedge strides 1, 2, 4, 8 and 16 gave
// 59.2, 60.5, 59.7, 61.5 and 65.7 ms per iteration.
int32_t sweep_publish_stride() {
    if (const char *e = std::getenv("PSST_SWEEP_PUBLISH_STRIDE")) {
        const long v = std::strtol(e, nullptr, 10);
        if (v > 0 && v <= 1024) return (int32_t)v;
    }
    return 1;
}

// Chunk schedule of the reference-ordered sweep with chunks of about
// chunk_cells cells: partition 1 = ranges of the device numbering (Morton),
// 2 = boxes of the per-axis potentials (no two chunks wait for each other),
// 0 = whichever gives the shorter modelled critical path with at most
// max_chunks chunks (0: no limit).  *used receives the partition taken.
psst::ChunkSchedule plan_sweep(const psst_mesh_desc *m, int32_t chunk_cells, int partition, int max_chunks,
                               const std::vector<int32_t> &lev, const int32_t *lnb, const std::vector<int32_t> &c2d,
                               const std::vector<int32_t> &d2c, int *used) {
    const int N = m->n_cells;
    psst::ChunkSchedule best;
    int64_t best_span = -1;
    for (int part : {1, 2}) {
        if (partition != 0 && partition != part) continue;
        std::vector<int32_t> chunk_of;
        int32_t n = 0;
        if (part == 1) {
            n = psst::morton_chunks(N, chunk_cells, c2d, chunk_of);
        } else {
            // boxes whose count exceeds the limit (the domain is not a whole number of
            // boxes) grow until it fits
            for (int32_t k = chunk_cells;; k = k + k / 8 + 1) {
                n = psst::potential_chunks(N, m->dim, k, m->list_offsets, m->list_faces, lnb, m->face_normal,
                                           m->n_faces, chunk_of);
                if (max_chunks <= 0 || n <= max_chunks || k >= N) break;
            }
        }
        if (partition == 0 && max_chunks > 0 && n > max_chunks) continue;
        // a publication stride above 1 can close a cycle of waits between chunks that
        // wait for each other; such a schedule publishes after every waited-for level
        const int32_t stride = sweep_publish_stride();
        psst::ChunkSchedule cs = psst::chunk_schedule(N, n, chunk_of, lev, m->list_offsets, lnb, c2d, d2c, stride);
        if (stride > 1 && !psst::schedule_completes(cs))
            cs = psst::chunk_schedule(N, n, chunk_of, lev, m->list_offsets, lnb, c2d, d2c, 1);
        const int64_t span = partition == 0 ? psst::schedule_stats(cs, kBlock).makespan_wait4 : 0;
        if (best_span < 0 || span < best_span) {
            best = std::move(cs);
            best_span = span;
            *used = part;
        }
    }
    if (best_span < 0) throw std::invalid_argument("no chunk partition within the workgroup limit");
    return best;
}

__global__ void k_canary(unsigned long long *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = 0x5053535443414e59ull;  // "PSSTCANY"
}

// Plain streaming copy of known traffic (counter calibration).
__global__ void __launch_bounds__(kBlock) k_copy(const double4 *__restrict__ a, double4 *__restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// The normalising residual row, by value.
struct Row5 {
    double v[5];
};

// Start of a chunk of a run: clears the residual slots and the control
// words, and loads the normalising row (or marks it as taken from the next
// row produced).
__global__ void __launch_bounds__(64) k_run_begin(RunScratch *rs, Row5 res0, int have_res0) {
    for (int i = (int)threadIdx.x; i < 5 * kResSlots; i += 64) rs->slots[i] = 0ull;
    if (threadIdx.x < 4) rs->ctl[threadIdx.x] = threadIdx.x == 0 ? -1 : 0;
    if (threadIdx.x < 5) rs->res0[threadIdx.x] = have_res0 ? res0.v[threadIdx.x] : -1.0;
}

// After every iteration of a run, one wave: the residual max-norm row from
// the slots (which it clears for the next iteration), appended to the
// chunk's rows; the drop max_j res_j / res0_j against the normalising row;
// and the stop when the drop meets the tolerance or the row is non-finite
// (psstFiniteVolumeGrid.cpp:369-383 logs the same row).  Once stopped, it
// and every kernel of the iterations queued behind return at once.
template <int WS>
__global__ void __launch_bounds__(64) k_run_check(RunScratch *rs, double tol) {
    if (rs->ctl[0] >= 0) return;
    const int lane = (int)threadIdx.x;
    double r[WS];
#pragma unroll
    for (int j = 0; j < WS; ++j) {
        unsigned long long v = rs->slots[j * kResSlots + lane];
        rs->slots[j * kResSlots + lane] = 0ull;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(v, off, 64);
            v = v > o ? v : o;
        }
        r[j] = __longlong_as_double((long long)v);
    }
    if (lane != 0) return;
    const int it = rs->ctl[1];
    bool bad = false;
#pragma unroll
    for (int j = 0; j < WS; ++j) {
        rs->rows[it * WS + j] = r[j];
        if (!(r[j] <= 1.0e300)) bad = true;  // NaN or inf
    }
    rs->ctl[1] = it + 1;
    if (rs->res0[0] < 0.0)
#pragma unroll
        for (int j = 0; j < WS; ++j) rs->res0[j] = r[j];
    const double d = residual_drop<WS>(r, rs->res0);
    if (bad) rs->ctl[2] = 1;
    if (bad || (tol > 0.0 && d <= tol)) rs->ctl[0] = it;
}

}  // namespace

// Workspace of psst_run_noalloc: one RunScratch per GPU, in a fixed device
// array (each GPU holds its own copy until the process ends).  In host code
// `scratch` names a host-side stand-in, never a device address: the device
// address of each GPU's copy is taken once per device with
// hipGetSymbolAddress (scratch_address) and only that address reaches
// kernels and copies.
constexpr size_t kScratchWords = (sizeof(RunScratch) + sizeof(unsigned long long) - 1) / sizeof(unsigned long long);
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
RunScratch *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `dev`, looked up on the first call for
// that device (with it current) and kept.
RunScratch *scratch_address(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw HipError("device index out of range for the fixed workspace");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) throw HipError("hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = static_cast<RunScratch *>(ptr);
    }
    return g_scratch_ptr[dev];
}

template <typename T>
void dispatch_iteration(psst_ctx *c, unsigned long long *res_row) {
    if (c->dim == 2) run_iteration<2, T>(c, res_row);
    else run_iteration<3, T>(c, res_row);
}

size_t state_bytes(const psst_ctx *c) { return c->s.precision == 32 ? sizeof(float) : sizeof(double); }

// canonical <-> device numbering of [rows][n] arrays (n = NT or N)
void to_device_order(const psst_ctx *c, const double *src, double *dst, int rows, int n) {
    for (int r = 0; r < rows; ++r)
        for (int d = 0; d < n; ++d) dst[(size_t)r * n + d] = src[(size_t)r * n + c->d2c[(size_t)d]];
}
void to_canonical_order(const psst_ctx *c, const double *src, double *dst, int rows, int n) {
    for (int r = 0; r < rows; ++r)
        for (int d = 0; d < n; ++d) dst[(size_t)r * n + c->d2c[(size_t)d]] = src[(size_t)r * n + d];
}

// device buffer (double or float) -> host double, device order
void download_raw(psst_ctx *c, const void *dev, double *out, size_t cnt) {
    if (c->s.precision == 64) {
        hip_check(hipMemcpyAsync(out, dev, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "D2H");
    } else {
        std::vector<float> f(cnt);
        hip_check(hipMemcpyAsync(f.data(), dev, cnt * sizeof(float), hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "D2H");
        for (size_t i = 0; i < cnt; ++i) out[i] = f[i];
    }
}

// device buffer (double or float) [rows][n] -> host double, canonical order
void download_canonical(psst_ctx *c, const void *dev, double *out, int rows, int n) {
    std::vector<double> tmp((size_t)rows * n);
    download_raw(c, dev, tmp.data(), tmp.size());
    to_canonical_order(c, tmp.data(), out, rows, n);
}

psst_status set_error(psst_ctx *c, const std::exception &e, psst_status st) {
    if (c) c->err = e.what();
    else g_create_error = e.what();
    return st;
}

// ctl and res0 of RunScratch, read back together after every chunk.
struct RunHead {
    int ctl[4];
    double res0[5];
};
static_assert(offsetof(RunScratch, res0) == offsetof(RunScratch, ctl) + sizeof(int) * 4, "RunScratch layout");

// Restores the context's stop pointer when a run ends (also on errors).
struct StopGuard {
    psst_ctx *c;
    ~StopGuard() { c->geom.stop = nullptr; }
};

// psst_run on the run control `rs`: chunks of at most `chunk_cap`
// iterations, each launched as begin, then per iteration the gradients, the
// RK stages and the check kernel; one read-back per chunk (control words,
// normalising row, and the rows straight into res_hist).  After the stop
// the state is that of the stopping iteration, with no snapshot or replay:
// the iterations queued behind it do nothing.
psst_status run_impl(psst_ctx *c, RunScratch *rs, int32_t n_iter, double tol, double *res_hist, int32_t *iters_done,
                     int chunk_cap) {
    if (!c->have_state) {
        c->err = "psst_run: no state (call psst_set_state first)";
        return PSST_BAD_ARG;
    }
    if (iters_done) *iters_done = 0;
    if (n_iter <= 0) return PSST_OK;
    if (c->batch) {
        c->err = "psst_run: the context of a batch runs through psst_batch_run";
        return PSST_BAD_ARG;
    }
    const int ws = c->dim + 2;
    try {
        StopGuard guard{c};
        c->geom.stop = &rs->ctl[0];
        const int chunk = std::max(1, std::min<int32_t>(n_iter, std::min(chunk_cap, kRunRows)));
        int done = 0;
        bool stopped = false, bad = false;
        while (done < n_iter && !stopped) {
            const int n = std::min(chunk, n_iter - done);
            Row5 r0{};
            for (int j = 0; j < ws; ++j) r0.v[j] = c->res0[j];
            launch(c, "run_begin", k_run_begin, dim3(1), dim3(64), rs, r0, (int)c->have_res0);
            const int cur0 = c->cur;
            for (int it = 0; it < n; ++it) {
                c->s.precision == 64 ? dispatch_iteration<double>(c, rs->slots) : dispatch_iteration<float>(c, rs->slots);
                if (ws == 4) launch(c, "run_check", k_run_check<4>, dim3(1), dim3(64), rs, tol);
                else launch(c, "run_check", k_run_check<5>, dim3(1), dim3(64), rs, tol);
            }
            RunHead h;
            hip_check(hipMemcpyAsync(&h, &rs->ctl[0], sizeof(h), hipMemcpyDeviceToHost, c->stream), "hipMemcpy D2H");
            if (res_hist)
                hip_check(hipMemcpyAsync(res_hist + (size_t)done * ws, rs->rows, sizeof(double) * (size_t)n * ws,
                                         hipMemcpyDeviceToHost, c->stream),
                          "hipMemcpy D2H");
            hip_check(hipStreamSynchronize(c->stream), "psst_run");
            if (c->timing) flush_timing(c);
            if (!c->have_res0) {  // the first row this context produced
                for (int j = 0; j < ws; ++j) c->res0[j] = h.res0[j];
                c->have_res0 = true;
            }
            if (h.ctl[3] != 0) {
                c->err = "reference-ordered sweep: a wait between chunks gave up (schedule or device fault)";
                return PSST_HIP_ERROR;
            }
            if (h.ctl[0] >= 0) {
                const int last = h.ctl[0];  // stopping row of this chunk
                c->cur = cur0 ^ (((last + 1) * c->s.rk_levels) & 1);
                if (res_hist && last + 1 < n)
                    std::memset(res_hist + (size_t)(done + last + 1) * ws, 0, sizeof(double) * (size_t)(n - last - 1) * ws);
                done += last + 1;
                stopped = true;
                bad = h.ctl[2] != 0;
            } else {
                done += n;
            }
            if (iters_done) *iters_done = done;
        }
        if (bad) {
            c->err = "non-finite residual: the state contains NaN/inf (a negative temperature or density upstream)";
            return PSST_NAN_DETECTED;
        }
        if (tol > 0.0 && !stopped) return PSST_NO_CONVERGENCE;
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char *psst_last_error(const psst_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

}  // extern "C"

// Context creation; seg (may be empty) lists the first interior cell of each
// independent instance of a batch.
static psst_status create_impl(const psst_mesh_desc *m, const psst_settings *s, const std::vector<int32_t> &seg,
                               psst_ctx **out) {
    if (!m || !s || !out) {
        g_create_error = "null argument";
        return PSST_BAD_ARG;
    }
    *out = nullptr;
    auto c = std::make_unique<psst_ctx>();
    try {
        if (m->dim != 2 && m->dim != 3) throw std::invalid_argument("dim must be 2 or 3");
        if (m->n_cells <= 0 || m->n_ghosts < 0 || m->n_faces <= 0) throw std::invalid_argument("empty mesh");
        if (s->rk_levels < 1 || s->rk_levels > 8) throw std::invalid_argument("rk_levels must be 1..8");
        if (s->flux < 1 || s->flux > 3) throw std::invalid_argument("flux must be 1 (AUSM), 2 (HLL) or 3 (HLLC)");
        if (s->order != 1 && s->order != 2) throw std::invalid_argument("order must be 1 or 2");
        if (s->order == 2 && (s->limiter < 1 || s->limiter > 4))
            throw std::invalid_argument("limiter must be 1..4 for second order");
        if (s->order == 2 && s->limiter == PSST_LIMITER_VENKATAKRISHNAN && !(s->venkat_k > 0.0))
            throw std::invalid_argument("venkat_k must be positive");
        if (s->time_mode < 0 || s->time_mode > 2) throw std::invalid_argument("time_mode must be 0, 1 or 2");
        if (s->precision != 64 && s->precision != 32) throw std::invalid_argument("precision must be 64 or 32");
        if (s->order == 2 && s->grad == PSST_GRAD_LEAST_SQUARES && (!m->lsq_r || !m->lsq_beta))
            throw std::invalid_argument("least squares needs lsq_r and lsq_beta");
        if (s->n_bc <= 0 || !s->bc) throw std::invalid_argument("boundary table missing");
        c->dim = m->dim;
        c->N = m->n_cells;
        c->G = m->n_ghosts;
        c->F = m->n_faces;
        c->NT = c->N + c->G;
        c->s = *s;
        c->bc_host.assign(s->bc, s->bc + s->n_bc);
        c->s.bc = nullptr;
        // Venkatakrishnan's eps^2 = (K h)^3, h = mean interior cell size
        // (sum of volumes in canonical order, then the dim-th root)
        double vsum = 0.0;
        for (int i = 0; i < m->n_cells; ++i) vsum += m->cell_volume[i];
        const double hbar = std::pow(vsum / m->n_cells, 1.0 / m->dim), kh = s->venkat_k * hbar;
        c->ph = Phys{s->gamma, s->rm, s->cp, s->cv, s->cfl, s->uniform_dt, kh * kh * kh};
        const int dim = c->dim, N = c->N, G = c->G, F = c->F, NT = c->NT;
        // ---- validation of the canonical topology (every index in range) ----
        const int32_t *f0 = m->face_cells, *f1 = m->face_cells + F;
        for (int f = 0; f < F; ++f) {
            if (f0[f] < 0 || f0[f] >= NT || f1[f] < 0 || f1[f] >= NT || f0[f] == f1[f])
                throw std::invalid_argument("face_cells out of range");
            if (f0[f] >= N && f1[f] >= N) throw std::invalid_argument("face between two ghosts");
        }
        if (m->visit_offsets[0] != 0 || m->list_offsets[0] != 0) throw std::invalid_argument("CSR offsets must start at 0");
        for (int i = 0; i < N; ++i) {
            if (m->visit_offsets[i + 1] < m->visit_offsets[i] || m->list_offsets[i + 1] < m->list_offsets[i])
                throw std::invalid_argument("CSR offsets not monotone");
        }
        c->nnz_visit = m->visit_offsets[N];
        c->nnz_list = m->list_offsets[N];
        std::vector<int> lnb_c((size_t)c->nnz_list);
        for (int i = 0; i < N; ++i) {
            for (int k = m->visit_offsets[i]; k < m->visit_offsets[i + 1]; ++k) {
                int f = m->visit_faces[k];
                if (f < 0 || f >= F || (f0[f] != i && f1[f] != i)) throw std::invalid_argument("visit list corrupt");
            }
            for (int k = m->list_offsets[i]; k < m->list_offsets[i + 1]; ++k) {
                int f = m->list_faces[k];
                if (f < 0 || f >= F || (f0[f] != i && f1[f] != i)) throw std::invalid_argument("face list corrupt");
                lnb_c[(size_t)k] = (f0[f] == i) ? f1[f] : f0[f];
            }
        }
        for (int i = 0; i < G; ++i) {
            int f = m->ghost_face[i];
            if (f < 0 || f >= F || (f0[f] != N + i && f1[f] != N + i)) throw std::invalid_argument("ghost_face corrupt");
            if (m->ghost_bc[i] < 0 || m->ghost_bc[i] >= s->n_bc) throw std::invalid_argument("ghost_bc out of range");
            const psst_bc &b = s->bc[m->ghost_bc[i]];
            int dk = b.dynamic_kind, tk = b.thermal_kind;
            if (dk != PSST_BC_SS_INLET && dk != PSST_BC_SS_OUTLET && dk != PSST_BC_SLIP_WALL && dk != PSST_BC_OUTLET)
                throw std::invalid_argument("unsupported dynamic boundary kind");
            if (tk != PSST_BCT_WALL_T && tk != PSST_BCT_WALL_Q && tk != PSST_BCT_INLET && tk != PSST_BCT_OUTLET)
                throw std::invalid_argument("unsupported thermal boundary kind");
        }
        // ---- device numbering: interior cells along a Morton curve, ghosts unchanged ----
        // typical cell spacing per axis: mean |dx_k| over interior faces whose
        // centre-to-centre vector is dominated by axis k
        double axis_h[3] = {0, 0, 0};
        axis_spacing(m, axis_h);
        psst::log_phase("context: mesh validated");
        std::vector<int32_t> d2c_int = psst::morton_order(dim, N, NT, m->cell_centre, seg, axis_h);
        psst::log_phase("context: cells renumbered");
        c->d2c.resize((size_t)NT);
        std::vector<int32_t> c2d((size_t)NT);
        for (int d = 0; d < N; ++d) c->d2c[(size_t)d] = d2c_int[(size_t)d];
        for (int d = N; d < NT; ++d) c->d2c[(size_t)d] = d;
        for (int d = 0; d < NT; ++d) c2d[(size_t)c->d2c[(size_t)d]] = d;
        const std::vector<int32_t> &d2c = c->d2c;
        std::vector<int32_t> f0d((size_t)F), f1d((size_t)F);
        for (int f = 0; f < F; ++f) {
            f0d[(size_t)f] = c2d[(size_t)f0[f]];
            f1d[(size_t)f] = c2d[(size_t)f1[f]];
        }
        std::vector<int32_t> voff((size_t)N + 1), vface((size_t)c->nnz_visit), loff((size_t)N + 1),
            lface((size_t)c->nnz_list), lnb((size_t)c->nnz_list), lsrc((size_t)c->nnz_list);
        voff[0] = loff[0] = 0;
        for (int d = 0; d < N; ++d) {
            const int cc = d2c[(size_t)d];
            int o = voff[(size_t)d];
            for (int k = m->visit_offsets[cc]; k < m->visit_offsets[cc + 1]; ++k) vface[(size_t)o++] = m->visit_faces[k];
            voff[(size_t)d + 1] = o;
            o = loff[(size_t)d];
            for (int k = m->list_offsets[cc]; k < m->list_offsets[cc + 1]; ++k) {
                lface[(size_t)o] = m->list_faces[k];
                lnb[(size_t)o] = c2d[(size_t)lnb_c[(size_t)k]];
                lsrc[(size_t)o] = k;
                ++o;
            }
            loff[(size_t)d + 1] = o;
        }
        std::vector<int> gnb((size_t)std::max(G, 1));
        for (int i = 0; i < G; ++i) {
            int f = m->ghost_face[i];
            gnb[(size_t)i] = (f0d[(size_t)f] == N + i) ? f1d[(size_t)f] : f0d[(size_t)f];
        }
        // neighbour 1 of face f in neighbour 0's frame (periodic faces)
        auto sh = [&](int k, int f) { return m->face_shift ? m->face_shift[(size_t)k * F + f] : 0.0; };
        auto x1 = [&](int k, int f) {
            const double x = m->cell_centre[(size_t)k * NT + f1[f]], t = sh(k, f);
            return t != 0.0 ? x + t : x;
        };
        std::vector<double> rlr3(3 * (size_t)F, 0.0);
        for (int f = 0; f < F; ++f)
            for (int k = 0; k < dim; ++k) rlr3[(size_t)k * F + f] = m->cell_centre[(size_t)k * NT + f0[f]] - x1(k, f);
        std::vector<double> vol_d((size_t)NT), cx_d(3 * (size_t)NT, 0.0), fn3(3 * (size_t)F, 0.0);
        for (int d = 0; d < NT; ++d) {
            vol_d[(size_t)d] = m->cell_volume[d2c[(size_t)d]];
            for (int k = 0; k < dim; ++k) cx_d[(size_t)k * NT + d] = m->cell_centre[(size_t)k * NT + d2c[(size_t)d]];
        }
        for (int k = 0; k < dim; ++k) std::memcpy(&fn3[(size_t)k * F], m->face_normal + (size_t)k * F, sizeof(double) * F);
        // projected areas for time step mode 2, face-list order (psstLocalTimeCalculation.cpp:192-202)
        std::vector<double> S(3 * (size_t)N, 0.0);
        for (int d = 0; d < N; ++d) {
            double sk[3] = {0, 0, 0};
            for (int k = loff[(size_t)d]; k < loff[(size_t)d + 1]; ++k) {
                int f = lface[(size_t)k];
                for (int x = 0; x < dim; ++x) sk[x] += m->face_area[f] * std::abs(m->face_normal[(size_t)x * F + f]);
            }
            for (int x = 0; x < 3; ++x) S[(size_t)x * N + d] = sk[x];
        }
        // least-squares weights w*theta_k per list entry (psstGradCalculation.cpp:309-356;
        // 3D: textbook QR weights), computed in canonical terms
        std::vector<double> lsqc(3 * (size_t)std::max(c->nnz_list, 1), 0.0);
        if (s->grad == PSST_GRAD_LEAST_SQUARES && m->lsq_r) {
            for (int d = 0; d < N; ++d) {
                const int i = d2c[(size_t)d];
                const double ra = m->lsq_r[(size_t)0 * N + i], rb = m->lsq_r[(size_t)1 * N + i],
                             rd = m->lsq_r[(size_t)2 * N + i], re = m->lsq_r[(size_t)4 * N + i],
                             rg = m->lsq_r[(size_t)5 * N + i], beta = m->lsq_beta[i];
                for (int k = loff[(size_t)d]; k < loff[(size_t)d + 1]; ++k) {
                    const int nb = lnb_c[(size_t)lsrc[(size_t)k]];
                    const int fl = lface[(size_t)k];
                    double diff[3] = {0, 0, 0}, w = 0.0, alpha[3] = {0, 0, 0}, theta[3] = {0, 0, 0};
                    for (int x = 0; x < dim; ++x) {
                        double xn = m->cell_centre[(size_t)x * NT + nb];
                        const double t = sh(x, fl);
                        if (t != 0.0) xn = (f0[fl] == i) ? xn + t : xn - t;
                        diff[x] = xn - m->cell_centre[(size_t)x * NT + i];
                        w += diff[x] * diff[x];
                    }
                    double ws = 1.0 / w;
                    if (dim == 2) {
                        alpha[0] = diff[0] / (ra * ra);
                        alpha[1] = (diff[1] - rb / ra * diff[0]) / (rd * rd);
                        theta[0] = alpha[0] - rb / ra * alpha[1];
                        theta[1] = alpha[1];
                    } else {
                        alpha[0] = diff[0] / (ra * ra);
                        alpha[1] = (diff[1] - rb / ra * diff[0]) / (rd * rd);
                        alpha[2] = (diff[2] - re / rd * diff[1] + beta * diff[0]) / (rg * rg);
                        theta[0] = alpha[0] - rb / ra * alpha[1] + beta * alpha[2];
                        theta[1] = alpha[1] - re / rd * alpha[2];
                        theta[2] = alpha[2];
                    }
                    for (int x = 0; x < 3; ++x) lsqc[(size_t)x * c->nnz_list + k] = ws * theta[x];
                }
            }
        }
        hip_check(hipStreamCreate(&c->stream), "hipStreamCreate");
        hip_check(hipGetDevice(&c->device), "hipGetDevice");
        // the fixed workspace of psst_run_noalloc on this device (its address
        // is taken here, once per device, with the device current)
        (void)scratch_address(c->device);
        {
            // a kernel of this library must demonstrably run: a HIP runtime
            // other than the one it was built for (e.g. one bundled with
            // another framework and loaded first under the same soname) can
            // leave launches without effect and without an error
            unsigned long long *probe = dev_alloc<unsigned long long>(c.get(), 1);
            unsigned long long zero = 0, got = 0;
            hip_check(hipMemcpy(probe, &zero, sizeof(zero), hipMemcpyHostToDevice), "hipMemcpy");
            hipLaunchKernelGGL(k_canary, dim3(1), dim3(64), 0, c->stream, probe);
            hip_check(hipGetLastError(), "canary launch");
            hip_check(hipStreamSynchronize(c->stream), "canary");
            hip_check(hipMemcpy(&got, probe, sizeof(got), hipMemcpyDeviceToHost), "hipMemcpy");
            if (got != 0x5053535443414e59ull) {
                int rv = 0;
                (void)hipRuntimeGetVersion(&rv);
                throw HipError("gfx950 kernels of libpsst did not run (HIP runtime version " + std::to_string(rv) +
                               " in this process): load libpsst before any library that bundles its own HIP "
                               "runtime, or use the system ROCm runtime");
            }
        }
        // reference-ordered schedule for time step mode 1: level of a cell =
        // 1 + max level of its interior face-list neighbours earlier in the
        // update loop (canonical index order); stored as device ids
        if (s->time_mode == PSST_DT_LOCAL1 && s->dt_ordering == 0) {
            const std::vector<int32_t> lev = psst::ordered_levels(N, m->list_offsets, lnb_c.data());
            int maxl = 0;
            for (int i = 0; i < N; ++i) maxl = std::max(maxl, lev[(size_t)i]);
            c->level_off.assign((size_t)maxl + 2, 0);
            for (int i = 0; i < N; ++i) c->level_off[(size_t)lev[(size_t)i] + 1]++;
            for (int l = 0; l <= maxl; ++l) c->level_off[(size_t)l + 1] += c->level_off[(size_t)l];
            std::vector<int> cells((size_t)N), fill(c->level_off.begin(), c->level_off.end() - 1);
            int widest = 0;
            for (int i = 0; i < N; ++i) cells[(size_t)fill[(size_t)lev[(size_t)i]]++] = c2d[(size_t)i];
            for (int l = 0; l <= maxl; ++l) widest = std::max(widest, c->level_off[(size_t)l + 1] - c->level_off[(size_t)l]);
            c->max_level = widest;
            c->level_cells = dev_upload(c.get(), cells.data(), cells.size());
            c->level_off_dev = dev_upload(c.get(), c->level_off.data(), c->level_off.size());
            // tables of the ordered sweep (single contexts; a batch keeps the
            // per-level launches, which it captures into graphs)
            int W = 0;
            for (int i = 0; i < N; ++i) W = std::max(W, m->list_offsets[i + 1] - m->list_offsets[i]);
            if (seg.empty() && W <= kOrderedW) {
                W = W <= 4 ? 4 : kOrderedW;  // the width the sweep kernels are built for
                // Large meshes with wide levels sweep by chunks (one workgroup
                // per chunk, waits only between chunks that share faces); the
                // tables are then in chunk order, else in level order.
                // workgroups of 128 threads when no level of a chunk is wider (twice
                // the chunks fit one cooperative launch), else 256
                const bool f64 = s->precision == 64;
                auto capacity_of = [&](int threads) {
                    return dim == 2 ? (f64 ? chunk_sweep_capacity<2, double>(W, threads)
                                           : chunk_sweep_capacity<2, float>(W, threads))
                                    : (f64 ? chunk_sweep_capacity<3, double>(W, threads)
                                           : chunk_sweep_capacity<3, float>(W, threads));
                };
                const bool large = !(widest <= 256 || N <= 65536);
                // one-wave workgroups (64 threads) where chunks of about 3,600 cells keep
                // every entry within 64 cells: an entry's barriers are then within one
                // wave, and more, smaller chunks fit one launch (60.6 -> 52-54 ms per
                // iteration on the 4.9M-cell wedge); else 128, else 256
                const int forced = sweep_threads();
                std::vector<int> sizes = forced ? std::vector<int>{forced} : std::vector<int>{64, 128, kBlock};
                int threads = 0;
                bool chunked = false;
                psst::ChunkSchedule cs;
                std::vector<int32_t> chunk_of_d;  // chunk of every device interior cell
                int used = 0;
                for (size_t si = 0; si < sizes.size() && large && NT < kRemote; ++si) {
                    const int cap = capacity_of(sizes[si]);
                    if (cap <= 0) continue;
                    psst::ChunkSchedule x = plan_sweep(m, sweep_chunk_cells(N, cap, sizes[si]), 0, cap, lev,
                                                       lnb_c.data(), c2d, d2c, &used);
                    int wide = 0;
                    for (size_t e = 0; e + 1 < x.ent_pos.size(); ++e) wide = std::max(wide, x.ent_pos[e + 1] - x.ent_pos[e]);
                    threads = sizes[si];
                    cs = std::move(x);
                    chunked = true;
                    if (wide <= threads) break;
                }
                if (chunked) {
                    // a chunk's entry tables must fit the kernel's LDS
                    for (int w = 0; w < cs.n_chunks && chunked; ++w) {
                        const int32_t e0 = cs.ch_ent[(size_t)w], e1 = cs.ch_ent[(size_t)w + 1];
                        chunked = e1 - e0 <= kChunkMaxEntries &&
                                  cs.ent_wait[(size_t)e1] - cs.ent_wait[(size_t)e0] <= kChunkMaxWaits;
                    }
                    const std::string msg =
                        !chunked ? std::string("context: chunk tables too large, ordered sweep by levels")
                                 : "context: ordered sweep by " + std::to_string(cs.n_chunks) +
                                       (used == 1 ? " Morton chunks" : " potential boxes") + " of " +
                                       std::to_string(threads) + " threads";
                    psst::log_phase(msg.c_str());
                }
                if (chunked) {
                    chunk_of_d.resize((size_t)N);
                    for (int w = 0; w < cs.n_chunks; ++w)
                        for (int p = cs.ent_pos[(size_t)cs.ch_ent[(size_t)w]];
                             p < cs.ent_pos[(size_t)cs.ch_ent[(size_t)w + 1]]; ++p)
                            chunk_of_d[(size_t)cs.pos_cell[(size_t)p]] = w;
                    std::vector<int2> waits(cs.wait_chunk.size());
                    for (size_t i = 0; i < waits.size(); ++i) waits[i] = make_int2(cs.wait_chunk[i], cs.wait_level[i]);
                    ChunkTables &ct = c->ct;
                    ct.ch_ent = dev_upload(c.get(), cs.ch_ent.data(), cs.ch_ent.size());
                    ct.ent_pos = dev_upload(c.get(), cs.ent_pos.data(), cs.ent_pos.size());
                    ct.ent_level = dev_upload(c.get(), cs.ent_level.data(), cs.ent_level.size());
                    ct.ent_wait = dev_upload(c.get(), cs.ent_wait.data(), cs.ent_wait.size());
                    ct.waits = dev_upload(c.get(), waits.data(), waits.size());
                    ct.progress = dev_alloc<int>(c.get(), (size_t)cs.n_chunks);
                    ct.timed_out = dev_alloc<int>(c.get(), 1);  // used outside psst_run's control words
                    hip_check(hipMemset(ct.timed_out, 0, sizeof(int)), "hipMemset");
                    ct.n_chunks = cs.n_chunks;
                    ct.threads = threads;
                }
                const std::vector<int32_t> &order = chunked ? cs.pos_cell : cells;
                std::vector<int> onb((size_t)W * N, -1);
                std::vector<double> ofn((size_t)dim * W * N, 0.0), ofa((size_t)W * N, 0.0), ovol((size_t)N);
                for (int p = 0; p < N; ++p) {
                    const int i = d2c[(size_t)order[(size_t)p]];
                    ovol[(size_t)p] = m->cell_volume[i];
                    int w = 0;
                    for (int k = m->list_offsets[i]; k < m->list_offsets[i + 1]; ++k, ++w) {
                        const int f = m->list_faces[k], nbc = lnb_c[(size_t)k];
                        int e = c2d[(size_t)nbc];
                        if (nbc < i) {
                            e |= kEarlier;
                            if (chunked && chunk_of_d[(size_t)c2d[(size_t)nbc]] != chunk_of_d[(size_t)order[(size_t)p]])
                                e |= kRemote;
                        }
                        onb[(size_t)w * N + p] = e;
                        ofa[(size_t)w * N + p] = m->face_area[f];
                        for (int x = 0; x < dim; ++x)
                            ofn[((size_t)x * W + w) * N + p] = m->face_normal[(size_t)x * F + f];
                    }
                }
                if (chunked) {
                    // a cell is a producer when a cell of another chunk comes later and reads it
                    std::vector<int32_t> pcell(cs.pos_cell);
                    std::vector<char> producer((size_t)N, 0);
                    for (int i = 0; i < N; ++i)
                        for (int k = m->list_offsets[i]; k < m->list_offsets[i + 1]; ++k) {
                            const int nbc = lnb_c[(size_t)k];
                            if (nbc < i && chunk_of_d[(size_t)c2d[(size_t)nbc]] != chunk_of_d[(size_t)c2d[(size_t)i]])
                                producer[(size_t)c2d[(size_t)nbc]] = 1;
                        }
                    for (int p = 0; p < N; ++p)
                        if (producer[(size_t)pcell[(size_t)p]]) pcell[(size_t)p] |= kProducer;
                    c->ot.cell = dev_upload(c.get(), pcell.data(), pcell.size());
                } else {
                    c->ot.cell = c->level_cells;
                }
                {
                    std::vector<int32_t> opos((size_t)N);
                    for (int p = 0; p < N; ++p) opos[(size_t)order[(size_t)p]] = p;
                    c->ot.pos = dev_upload(c.get(), opos.data(), opos.size());
                }
                c->ot.vol = dev_upload(c.get(), ovol.data(), ovol.size());
                c->ot.nb = dev_upload(c.get(), onb.data(), onb.size());
                c->ot.fn = dev_upload(c.get(), ofn.data(), ofn.size());
                c->ot.fa = dev_upload(c.get(), ofa.data(), ofa.size());
                c->ot.R = dev_alloc<double>(c.get(), (size_t)(dim + 2) * N);
                c->ot.W = W;
            }
        }
        // ---- device geometry ----
        // Device face order: faces sorted by their lower device cell id, so
        // that the face-indexed kernels (unfused fluxes, Green-Gauss and
        // least-squares sweeps, the reference-ordered time step) read
        // neighbouring cells in neighbouring lanes.  Every order-dependent
        // sum walks a per-cell face list, which keeps the reference's order
        // whatever the face numbering.  The tile plan works on the canonical
        // host arrays and is not affected.
        std::vector<int32_t> forder((size_t)F), fpos((size_t)F);
        for (int f = 0; f < F; ++f) forder[(size_t)f] = f;
        std::stable_sort(forder.begin(), forder.end(), [&](int32_t a, int32_t b) {
            return std::min(f0d[(size_t)a], f1d[(size_t)a]) < std::min(f0d[(size_t)b], f1d[(size_t)b]);
        });
        for (int p = 0; p < F; ++p) fpos[(size_t)forder[(size_t)p]] = p;
        // [rows][F] arrays in device face order, and face ids as device positions
        auto upload_by_face = [&](const auto *src, size_t rows) {
            using V = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
            std::vector<V> out(rows * (size_t)F);
            for (size_t r = 0; r < rows; ++r)
                for (int p = 0; p < F; ++p) out[r * F + p] = src[r * F + (size_t)forder[(size_t)p]];
            return dev_upload(c.get(), out.data(), out.size());
        };
        auto upload_positions = [&](const int32_t *ids, size_t n) {
            std::vector<int32_t> out(n);
            for (size_t i = 0; i < n; ++i) out[i] = fpos[(size_t)ids[i]];
            return dev_upload(c.get(), out.data(), out.size());
        };
        Geom &g = c->geom;
        g.N = N;
        g.G = G;
        g.NT = NT;
        g.F = F;
        g.vol = dev_upload(c.get(), vol_d.data(), vol_d.size());
        g.cx = dev_upload(c.get(), cx_d.data(), cx_d.size());
        g.fn = upload_by_face(fn3.data(), 3);
        g.fa = upload_by_face(m->face_area, 1);
        g.beta = upload_by_face(m->face_beta, 1);
        g.f0 = upload_by_face(f0d.data(), 1);
        g.f1 = upload_by_face(f1d.data(), 1);
        g.rlr = upload_by_face(rlr3.data(), 3);
        g.vface = upload_positions(vface.data(), vface.size());
        g.lface = upload_positions(lface.data(), lface.size());
        g.gface = upload_positions(m->ghost_face, (size_t)G);
        g.voff = dev_upload(c.get(), voff.data(), voff.size());
        g.loff = dev_upload(c.get(), loff.data(), loff.size());
        g.lnb = dev_upload(c.get(), lnb.data(), lnb.size());
        g.lsqc = dev_upload(c.get(), lsqc.data(), lsqc.size());
        g.S = dev_upload(c.get(), S.data(), S.size());
        g.gnb = dev_upload(c.get(), gnb.data(), (size_t)G);
        g.gbc = dev_upload(c.get(), m->ghost_bc, (size_t)G);
        {
            std::vector<int32_t> rank(d2c.begin(), d2c.end());
            g.rank = dev_upload(c.get(), rank.data(), rank.size());
        }
        std::vector<BCDev> bcd((size_t)s->n_bc);
        for (int i = 0; i < s->n_bc; ++i) {
            bcd[(size_t)i].dyn = s->bc[i].dynamic_kind;
            bcd[(size_t)i].th = s->bc[i].thermal_kind;
            for (int d = 0; d < 3; ++d) bcd[(size_t)i].u[d] = s->bc[i].velocity[d];
            bcd[(size_t)i].p = s->bc[i].pressure;
            bcd[(size_t)i].T = s->bc[i].temperature;
        }
        c->bc = dev_upload(c.get(), bcd.data(), bcd.size());
        const size_t sb = state_bytes(c.get());
        // ---- tile plan of the fused kernels ----
        const bool ordered_dt1 = (s->time_mode == PSST_DT_LOCAL1 && s->dt_ordering == 0);
        if (s->fused) {
            // Tile size: the caller's, or the first of a list whose stage
            // kernel LDS lets a CU (160 KB) hold 16 or more waves.  The lists
            // start at what was fastest on the 3840x1280 wedge (FP64 224 cells
            // / 256 threads, FP32 768 / 512); a 3D hexahedral tile carries
            // about as many halo cells as own cells, and there 128-cell FP64
            // tiles (four workgroups per CU) ran 31% faster than 224-cell ones
            // (two).
            if (s->tile_threads != 0 && s->tile_threads != 256 && s->tile_threads != 512)
                throw std::invalid_argument("tile_threads must be 256 or 512");
            std::vector<int> cands;
            if (s->tile_cells > 0) {
                if (s->tile_cells < 32 || s->tile_cells > 4096) throw std::invalid_argument("tile_cells must be 32..4096");
                cands = {s->tile_cells};
            } else if (s->precision == 64) {
                cands = {224, 192, 160, 128, 96};
            } else {
                cands = {768, 512, 384, 256, 192, 128};
            }
            const double cost_slot = (double)(dim + 3) * sb + sizeof(int), cost_face = (double)(dim + 2) * sb;
            // 2D: a tile's faces grouped by the axis their normal is closest
            // to, boundary faces last.  The lanes of a wavefront of the
            // gradient kernel then read the same face of neighbouring cells
            // at neighbouring addresses.  On the 4.9M-cell wedge: Green-Gauss
            // 0.235 vs 0.243 ms (FP64), 0.158 vs 0.176 ms (FP32); the stage
            // kernel 0.372 vs 0.370 ms (FP64), 0.256 vs 0.258 ms (FP32); all
            // bitwise unchanged.  On the 12.6M-hexahedra wedge the 3D stage
            // kernel took 2.40 vs 2.05 ms, so 3D keeps the order of first
            // appearance over the tile's cells.
            std::vector<uint8_t> fclass(dim == 2 ? (size_t)F : 0);
            for (int f = 0; f < (dim == 2 ? F : 0); ++f) {
                if (f0d[(size_t)f] >= N || f1d[(size_t)f] >= N) {
                    fclass[(size_t)f] = (uint8_t)dim;
                    continue;
                }
                int best = 0;
                for (int k = 1; k < dim; ++k)
                    if (std::fabs(m->face_normal[(size_t)k * F + f]) > std::fabs(m->face_normal[(size_t)best * F + f]))
                        best = k;
                fclass[(size_t)f] = (uint8_t)best;
            }
            psst::TilePlan tp;
            size_t shm_stage = 0, shm_grad = 0;
            for (size_t ci = 0; ci < cands.size(); ++ci) {
                tp = psst::plan_tiles(N, G, F, cands[ci], f0d.data(), f1d.data(), voff.data(), vface.data(),
                                      loff.data(), lface.data(), seg, cost_slot, cost_face,
                                      fclass.empty() ? nullptr : fclass.data());
                c->max_slots = tp.max_slots;
                c->max_faces = tp.max_faces;
                shm_stage = stage_shm(c.get(), sb);
                shm_grad = grad_shm(c.get(), sb);
                psst::log_phase(("context: tile plan of " + std::to_string(cands[ci]) + " cells, stage LDS " +
                                 std::to_string(shm_stage) + " B").c_str());
                const bool last = ci + 1 == cands.size();
                const int thr = s->tile_threads ? s->tile_threads : (cands[ci] >= 512 ? 512 : 256);
                const size_t waves = (160 * 1024 / std::max<size_t>(shm_stage, 1)) * (size_t)(thr / 64);
                if (waves >= 16 || (last && shm_stage <= 160 * 1024)) break;
            }
            const int tile = tp.tile;
            c->tile_threads = s->tile_threads ? s->tile_threads : (tile >= 512 ? 512 : 256);
            c->tile_cells = tile;
            c->tile_cell_begin = tp.cell_begin;
            if (shm_stage <= 160 * 1024) {
                const int TF = (int)tp.face_id.size();
                std::vector<double> tfn((size_t)dim * TF), tfa((size_t)TF), trlr((size_t)dim * TF), tbeta((size_t)TF);
                for (int j = 0; j < TF; ++j) {
                    const int f = tp.face_id[(size_t)j];
                    for (int k = 0; k < dim; ++k) {
                        tfn[(size_t)k * TF + j] = m->face_normal[(size_t)k * F + f];
                        trlr[(size_t)k * TF + j] = rlr3[(size_t)k * F + f];
                    }
                    tfa[(size_t)j] = m->face_area[f];
                    tbeta[(size_t)j] = m->face_beta[f];
                }
                std::vector<double> tlsq((size_t)dim * std::max(tp.W_list, 1) * N, 0.0);
                if (s->grad == PSST_GRAD_LEAST_SQUARES)
                    for (int d = 0; d < N; ++d)
                        for (int k = loff[(size_t)d], w = 0; k < loff[(size_t)d + 1]; ++k, ++w)
                            for (int x = 0; x < dim; ++x)
                                tlsq[((size_t)x * tp.W_list + w) * N + d] = lsqc[(size_t)x * c->nnz_list + k];
                TileDev &td = c->td;
                td.cell_begin = dev_upload(c.get(), tp.cell_begin.data(), tp.cell_begin.size());
                td.halo_off = dev_upload(c.get(), tp.halo_off.data(), tp.halo_off.size());
                td.halo = dev_upload(c.get(), tp.halo.data(), tp.halo.size());
                td.face_off = dev_upload(c.get(), tp.face_off.data(), tp.face_off.size());
                td.ends = dev_upload(c.get(), tp.face_ends.data(), tp.face_ends.size());
                td.fn = dev_upload(c.get(), tfn.data(), tfn.size());
                td.fa = dev_upload(c.get(), tfa.data(), tfa.size());
                td.rlr = dev_upload(c.get(), trlr.data(), trlr.size());
                td.beta = dev_upload(c.get(), tbeta.data(), tbeta.size());
                if (dim == 2) {
                    std::vector<double4> tgg((size_t)TF);
                    for (int j = 0; j < TF; ++j)
                        tgg[(size_t)j] = make_double4(tfa[(size_t)j], tbeta[(size_t)j], tfn[(size_t)j], tfn[(size_t)TF + j]);
                    td.gg = dev_upload(c.get(), tgg.data(), tgg.size());
                    std::vector<double4> tsf((size_t)TF);
                    for (int j = 0; j < TF; ++j)
                        tsf[(size_t)j] = make_double4(tfn[(size_t)j], tfn[(size_t)TF + j], trlr[(size_t)j], trlr[(size_t)TF + j]);
                    td.sf = dev_upload(c.get(), tsf.data(), tsf.size());
                } else {
                    std::vector<double2> tsf3((size_t)3 * TF);
                    for (int j = 0; j < TF; ++j) {
                        tsf3[(size_t)3 * j] = make_double2(tfn[(size_t)j], tfn[(size_t)TF + j]);
                        tsf3[(size_t)3 * j + 1] = make_double2(tfn[(size_t)2 * TF + j], trlr[(size_t)j]);
                        tsf3[(size_t)3 * j + 2] = make_double2(trlr[(size_t)TF + j], trlr[(size_t)2 * TF + j]);
                    }
                    td.sf3 = dev_upload(c.get(), tsf3.data(), tsf3.size());
                }
                td.ghost_off = dev_upload(c.get(), tp.ghost_off.data(), tp.ghost_off.size());
                td.ghost = dev_upload(c.get(), tp.ghost.data(), tp.ghost.size());
                td.visit = reinterpret_cast<const short *>(dev_upload(c.get(), tp.visit_ell.data(), tp.visit_ell.size()));
                td.list = reinterpret_cast<const short *>(dev_upload(c.get(), tp.list_ell.data(), tp.list_ell.size()));
                td.lslot =
                    reinterpret_cast<const short *>(dev_upload(c.get(), tp.list_slot.data(), tp.list_slot.size()));
                td.lsqc = dev_upload(c.get(), tlsq.data(), tlsq.size());
                td.TF = TF;
                td.Wv = tp.W_visit;
                td.Wl = tp.W_list;
                td.MS = tp.max_slots;
                td.MF = tp.max_faces;
                td.n_tiles = tp.n_tiles;
                c->n_tiles = tp.n_tiles;
                c->shm_stage = shm_stage;
                c->shm_grad = shm_grad;
                c->tiled = true;
                // the reference-ordered time step keeps its sweep; the tiles
                // serve its gradients and residual pass
                c->fused = !ordered_dt1;
            }
        }
        psst::log_phase("context: geometry and tiles uploaded");
        // ---- state: sized for FP64 whatever the precision, so that
        // psst_set_precision can switch in place ----
        const size_t sb8 = sizeof(double);
        for (int i = 0; i < 2; ++i) {
            hip_check(hipMalloc(&c->q[i], sb8 * (size_t)(dim + 3) * NT), "hipMalloc");
            c->allocs.push_back(c->q[i]);
        }
        hip_check(hipMalloc(&c->W0, sb8 * (size_t)(dim + 2) * N), "hipMalloc");
        c->allocs.push_back(c->W0);
        hip_check(hipMalloc(&c->grad, sb8 * (size_t)(dim == 2 ? GradLayout<2>::STRIDE : GradLayout<3>::STRIDE) * N),
                  "hipMalloc");
        c->allocs.push_back(c->grad);
        hip_check(hipMalloc(&c->FA, sb8 * (size_t)(dim + 2) * F), "hipMalloc");
        c->allocs.push_back(c->FA);
        c->run = dev_alloc<RunScratch>(c.get(), 1);
        c->geom.stop = nullptr;
        c->kernel_names = {"grad_gg",      "grad_lsq",
                           "flux",         "update",
                           "update_level", "update_levels_block",
                           "bc",
                           "stage",        "grad_gg_tile",
                           "grad_lsq_tile", "batch_check",
                           "ordered_residual", "update_levels_wg",
                           "update_chunks", "stage_residual"};
    } catch (const std::exception &e) {
        g_create_error = e.what();
        for (void *p : c->allocs) (void)hipFree(p);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        return dynamic_cast<const HipError *>(&e) ? PSST_HIP_ERROR : PSST_BAD_ARG;
    }
    *out = c.release();
    return PSST_OK;
}

extern "C" {

psst_status psst_ctx_create(const psst_mesh_desc *m, const psst_settings *s, psst_ctx **out) {
    return create_impl(m, s, {}, out);
}

psst_status psst_debug_sweep_schedule(const psst_mesh_desc *m, int32_t chunk_cells, int32_t partition,
                                      int64_t *stats) {
    if (!m || !stats || m->n_cells <= 0 || m->n_faces <= 0 || (m->dim != 2 && m->dim != 3) || partition < 0 ||
        partition > 2) {
        g_create_error = "psst_debug_sweep_schedule: bad argument";
        return PSST_BAD_ARG;
    }
    try {
        const int N = m->n_cells, F = m->n_faces, NT = m->n_cells + m->n_ghosts;
        const int32_t *f0 = m->face_cells, *f1 = m->face_cells + F;
        const int nnz = m->list_offsets[N];
        std::vector<int32_t> lnb((size_t)std::max(nnz, 1));
        for (int i = 0; i < N; ++i)
            for (int k = m->list_offsets[i]; k < m->list_offsets[i + 1]; ++k) {
                const int f = m->list_faces[k];
                if (f < 0 || f >= F || (f0[f] != i && f1[f] != i)) throw std::invalid_argument("face list corrupt");
                lnb[(size_t)k] = f0[f] == i ? f1[f] : f0[f];
            }
        double axis_h[3];
        axis_spacing(m, axis_h);
        const std::vector<int32_t> d2c = psst::morton_order(m->dim, N, NT, m->cell_centre, {}, axis_h);
        std::vector<int32_t> c2d((size_t)N);
        for (int d = 0; d < N; ++d) c2d[(size_t)d2c[(size_t)d]] = d;
        const std::vector<int32_t> lev = psst::ordered_levels(N, m->list_offsets, lnb.data());
        const int32_t k = chunk_cells > 0 ? chunk_cells : sweep_chunk_cells(N, 0, 128);
        int used = 0;
        const psst::ChunkSchedule cs = plan_sweep(m, k, partition, 0, lev, lnb.data(), c2d, d2c, &used);
        std::vector<int32_t> chunk_of((size_t)N);
        for (int w = 0; w < cs.n_chunks; ++w)
            for (int p = cs.ent_pos[(size_t)cs.ch_ent[(size_t)w]]; p < cs.ent_pos[(size_t)cs.ch_ent[(size_t)w + 1]]; ++p)
                chunk_of[(size_t)d2c[(size_t)cs.pos_cell[(size_t)p]]] = w;
        const psst::ScheduleStats st =
            psst::check_schedule(cs, N, chunk_of, lev, m->list_offsets, lnb.data(), c2d, d2c, kBlock);
        const int64_t out[10] = {st.levels,       st.chunks,        st.entries,        st.waits,   st.publishes,
                                 st.widest_entry, st.makespan_free, st.makespan_wait4, (int64_t)k, (int64_t)used};
        std::memcpy(stats, out, sizeof(out));
        return PSST_OK;
    } catch (const std::exception &e) {
        g_create_error = e.what();
        return PSST_BAD_ARG;
    }
}

psst_status psst_set_state(psst_ctx *c, const double *prim) {
    if (!c || !prim) return PSST_BAD_ARG;
    try {
        const size_t n = (size_t)(c->dim + 3) * c->NT;
        std::vector<double> dev_prim(n);
        to_device_order(c, prim, dev_prim.data(), c->dim + 3, c->NT);
        if (c->s.precision == 64) {
            hip_check(hipMemcpy(c->q[0], dev_prim.data(), n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
            hip_check(hipMemcpy(c->q[1], dev_prim.data(), n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        } else {
            double *tmp = nullptr;
            hip_check(hipMalloc(&tmp, n * sizeof(double)), "hipMalloc");
            hip_check(hipMemcpy(tmp, dev_prim.data(), n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
            hipLaunchKernelGGL((k_convert<double, float>), dim3(blocks((long)n)), dim3(kBlock), 0, c->stream, tmp,
                               (float *)c->q[0], n);
            hipLaunchKernelGGL((k_convert<double, float>), dim3(blocks((long)n)), dim3(kBlock), 0, c->stream, tmp,
                               (float *)c->q[1], n);
            hip_check(hipStreamSynchronize(c->stream), "convert");
            hip_check(hipFree(tmp), "hipFree");
        }
        c->cur = 0;
        if (c->dim == 2) {
            c->s.precision == 64 ? run_set_w<2, double>(c, 0, c->N) : run_set_w<2, float>(c, 0, c->N);
        } else {
            c->s.precision == 64 ? run_set_w<3, double>(c, 0, c->N) : run_set_w<3, float>(c, 0, c->N);
        }
        hip_check(hipStreamSynchronize(c->stream), "set_w");
        c->have_state = true;
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

psst_status psst_set_residual_reference(psst_ctx *c, const double *res0) {
    if (!c || !res0) return PSST_BAD_ARG;
    for (int j = 0; j < c->dim + 2; ++j) c->res0[j] = res0[j];
    c->have_res0 = true;
    return PSST_OK;
}

psst_status psst_run(psst_ctx *c, int32_t n_iter, double tol, double *res_hist, int32_t *iters_done) {
    if (!c) return PSST_BAD_ARG;
    // chunks of at most 64 iterations with a tolerance (after the stop the
    // rest of a chunk's queued iterations return at once), else of the
    // run rows' capacity
    return run_impl(c, c->run, n_iter, tol, res_hist, iters_done, tol > 0.0 ? 64 : kRunRows);
}

psst_status psst_run_noalloc(psst_ctx *c, int32_t n_iter, double tol, double *res_hist, int32_t *iters_done) {
    if (!c) return PSST_BAD_ARG;
    if (n_iter > kRunRows) {
        c->err = "psst_run_noalloc: n_iter exceeds the fixed workspace's 1024 residual rows per call";
        return PSST_BAD_ARG;
    }
    if (c->timing) {
        c->err = "psst_run_noalloc: kernel timing creates events; disable it (psst_kernel_timing(ctx, 0))";
        return PSST_BAD_ARG;
    }
    RunScratch *rs = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        rs = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
    }
    if (!rs) {
        c->err = "psst_run_noalloc: no workspace address for the context's device";
        return PSST_HIP_ERROR;
    }
    // one chunk: the whole call fits the workspace's rows
    return run_impl(c, rs, n_iter, tol, res_hist, iters_done, kRunRows);
}

psst_status psst_get_state(psst_ctx *c, double *prim) {
    if (!c || !prim) return PSST_BAD_ARG;
    try {
        download_canonical(c, c->q[c->cur], prim, c->dim + 3, c->NT);
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

psst_status psst_get_conservative(psst_ctx *c, double *w) {
    if (!c || !w) return PSST_BAD_ARG;
    try {
        download_canonical(c, c->W0, w, c->dim + 2, c->N);
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

psst_status psst_debug_gradients(psst_ctx *c, double *grads) {
    if (!c || !grads) return PSST_BAD_ARG;
    try {
        if (c->dim == 2) {
            c->s.precision == 64 ? run_grad<2, double>(c) : run_grad<2, float>(c);
        } else {
            c->s.precision == 64 ? run_grad<3, double>(c) : run_grad<3, float>(c);
        }
        {
            const int nc = (c->dim + 2) * c->dim;
            const int stride = c->dim == 2 ? GradLayout<2>::STRIDE : GradLayout<3>::STRIDE;
            std::vector<double> aos((size_t)stride * c->N), soa((size_t)nc * c->N);
            download_raw(c, c->grad, aos.data(), aos.size());
            for (int d = 0; d < c->N; ++d)
                for (int v = 0; v < nc; ++v) soa[(size_t)v * c->N + d] = aos[(size_t)d * stride + v];
            to_canonical_order(c, soa.data(), grads, nc, c->N);
        }
        if (c->timing) flush_timing(c);
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

psst_status psst_debug_residual(psst_ctx *c, double *residual) {
    if (!c || !residual) return PSST_BAD_ARG;
    try {
        double *out = nullptr;
        const size_t n = (size_t)(c->dim + 2) * c->N;
        hip_check(hipMalloc(&out, n * sizeof(double)), "hipMalloc");
        if (c->dim == 2) {
            c->s.precision == 64 ? run_debug_residual<2, double>(c, out) : run_debug_residual<2, float>(c, out);
        } else {
            c->s.precision == 64 ? run_debug_residual<3, double>(c, out) : run_debug_residual<3, float>(c, out);
        }
        std::vector<double> tmp(n);
        hip_check(hipMemcpyAsync(tmp.data(), out, n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "D2H");
        hip_check(hipFree(out), "hipFree");
        to_canonical_order(c, tmp.data(), residual, c->dim + 2, c->N);
        if (c->timing) flush_timing(c);
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

psst_status psst_kernel_timing(psst_ctx *c, int32_t enable) {
    if (!c) return PSST_BAD_ARG;
    try {
        if (!enable) flush_timing(c);
        c->timing = enable != 0;
        if (enable) c->stats.clear();
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

psst_status psst_kernel_stats(psst_ctx *c, const char *name, int64_t *count, double *mean_ms) {
    if (!c || !name) return PSST_BAD_ARG;
    try {
        flush_timing(c);
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
    auto it = c->stats.find(name);
    if (count) *count = it == c->stats.end() ? 0 : it->second.count;
    if (mean_ms) *mean_ms = (it == c->stats.end() || it->second.count == 0) ? 0.0 : it->second.total_ms / it->second.count;
    return PSST_OK;
}

psst_status psst_kernel_names(psst_ctx *c, char *buf, int32_t len) {
    if (!c || !buf || len <= 0) return PSST_BAD_ARG;
    std::string s;
    for (auto &n : c->kernel_names) {
        if (!s.empty()) s += ';';
        s += n;
    }
    if ((int32_t)s.size() + 1 > len) return PSST_BAD_ARG;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return PSST_OK;
}

psst_status psst_copy_calibration(int64_t bytes, int32_t reps, double *mean_ms) {
    if (bytes <= 0 || reps <= 0) return PSST_BAD_ARG;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    psst_status rc = PSST_OK;
    try {
        const size_t n = (size_t)bytes / sizeof(double4);
        hip_check(hipMalloc(&a, n * sizeof(double4)), "hipMalloc");
        hip_check(hipMalloc(&b, n * sizeof(double4)), "hipMalloc");
        hip_check(hipMemset(a, 0, n * sizeof(double4)), "hipMemset");
        hip_check(hipStreamCreate(&st), "hipStreamCreate");
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        const unsigned grid = 256 * 8 * 4;
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, st, (const double4 *)a, (double4 *)b, n);
        hip_check(hipEventRecord(e0, st), "hipEventRecord");
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, st, (const double4 *)a, (double4 *)b, n);
        hip_check(hipGetLastError(), "k_copy");
        hip_check(hipEventRecord(e1, st), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (mean_ms) *mean_ms = ms / reps;
    } catch (const std::exception &e) {
        g_create_error = e.what();
        rc = PSST_HIP_ERROR;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

void psst_ctx_destroy(psst_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->batch) {
        for (auto &kv : c->batch->exec)
            for (hipGraphExec_t &e : kv.second)
                if (e) (void)hipGraphExecDestroy(e);
        if (c->batch->h_flags) (void)hipHostFree(c->batch->h_flags);
        if (c->batch->h_ring) (void)hipHostFree(c->batch->h_ring);
        if (c->batch->h_drop) (void)hipHostFree(c->batch->h_drop);
    }
    for (auto &p : c->pending) {
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"

// ===========================================================================
// Precision switch, batches of instances, autotuner
// ===========================================================================
namespace {

void destroy_graphs(psst_ctx *c) {
    if (!c->batch) return;
    for (auto &kv : c->batch->exec)
        for (hipGraphExec_t &e : kv.second)
            if (e) (void)hipGraphExecDestroy(e);
    c->batch->exec.clear();
}

BatchCheck batch_check_args(psst_ctx *c) {
    BatchState &b = *c->batch;
    BatchCheck k;
    k.n = b.n;
    k.H = b.H;
    k.res = b.res;
    k.res0 = b.res0;
    k.hist = b.hist;
    k.drop = b.drop;
    k.active = b.active;
    k.conv = b.conv;
    k.frozen_buf = b.frozen_buf;
    k.iter = b.iter;
    k.tile_begin = b.tile_begin_dev;
    k.list = b.list;
    k.count = b.count;
    return k;
}

void launch_check(psst_ctx *c, double tol, int cur_after, int rebuild_only) {
    const BatchCheck k = batch_check_args(c);
    if (c->dim == 2)
        launch(c, "batch_check", k_batch_check<4>, dim3(1), dim3(kCheckThreads), k, tol, cur_after, rebuild_only);
    else
        launch(c, "batch_check", k_batch_check<5>, dim3(1), dim3(kCheckThreads), k, tol, cur_after, rebuild_only);
}

// One batch iteration: gradients, the RK stages over the active tiles, and
// the per-instance bookkeeping.
void batch_iteration(psst_ctx *c, double tol) {
    c->s.precision == 64 ? dispatch_iteration<double>(c, nullptr) : dispatch_iteration<float>(c, nullptr);
    launch_check(c, tol, c->cur, 0);
}

// The same iteration as a captured graph (one per parity of the state
// buffers at its start: with an odd stage count the buffers alternate).
void batch_graph_iteration(psst_ctx *c, double tol) {
    BatchState &b = *c->batch;
    if (b.graph_precision != c->s.precision || b.graph_tol != tol) {
        destroy_graphs(c);
        b.graph_precision = c->s.precision;
        b.graph_tol = tol;
    }
    const int parity = c->cur;
    hipGraphExec_t &ex = b.exec[b.grid_tiles][parity];
    if (!ex) {
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
        try {
            batch_iteration(c, tol);
        } catch (...) {
            (void)hipStreamEndCapture(c->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            c->cur = parity;
            throw;
        }
        hip_check(hipStreamEndCapture(c->stream, &g), "hipStreamEndCapture");
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) ex = nullptr;
        hip_check(e, "hipGraphInstantiate");
        c->cur = parity;
    }
    hip_check(hipGraphLaunch(ex, c->stream), "hipGraphLaunch");
    c->cur = parity ^ (c->s.rk_levels & 1);
}

// Brings every instance's latest state into the current state buffer.
void batch_gather_states(psst_ctx *c) {
    BatchState &b = *c->batch;
    const size_t sb = state_bytes(c);
    for (int i = 0; i < b.n; ++i) {
        const int from = b.buf_of[(size_t)i];
        if (from == c->cur) continue;
        const char *src = static_cast<const char *>(c->q[from]);
        char *dst = static_cast<char *>(c->q[c->cur]);
        for (int v = 0; v < c->dim + 3; ++v) {
            const size_t ci = (size_t)v * c->NT + b.cell_begin[(size_t)i];
            const size_t nc = (size_t)(b.cell_begin[(size_t)i + 1] - b.cell_begin[(size_t)i]);
            hip_check(hipMemcpyAsync(dst + ci * sb, src + ci * sb, nc * sb, hipMemcpyDeviceToDevice, c->stream), "D2D");
            const size_t gi = (size_t)v * c->NT + b.N_tot + b.ghost_begin[(size_t)i];
            const size_t ng = (size_t)(b.ghost_begin[(size_t)i + 1] - b.ghost_begin[(size_t)i]);
            if (ng)
                hip_check(hipMemcpyAsync(dst + gi * sb, src + gi * sb, ng * sb, hipMemcpyDeviceToDevice, c->stream),
                          "D2D");
        }
        b.buf_of[(size_t)i] = c->cur;
    }
}

void upload_values(psst_ctx *c, void *dev_base, size_t offset, const std::vector<double> &vals) {
    if (vals.empty()) return;
    if (c->s.precision == 64) {
        hip_check(hipMemcpy(static_cast<double *>(dev_base) + offset, vals.data(), vals.size() * sizeof(double),
                            hipMemcpyHostToDevice),
                  "hipMemcpy");
    } else {
        std::vector<float> f(vals.begin(), vals.end());
        hip_check(hipMemcpy(static_cast<float *>(dev_base) + offset, f.data(), f.size() * sizeof(float),
                            hipMemcpyHostToDevice),
                  "hipMemcpy");
    }
}

std::vector<double> download_values(psst_ctx *c, const void *dev_base, size_t offset, size_t n) {
    std::vector<double> out(n);
    if (!n) return out;
    if (c->s.precision == 64) {
        hip_check(hipMemcpy(out.data(), static_cast<const double *>(dev_base) + offset, n * sizeof(double),
                            hipMemcpyDeviceToHost),
                  "hipMemcpy");
    } else {
        std::vector<float> f(n);
        hip_check(hipMemcpy(f.data(), static_cast<const float *>(dev_base) + offset, n * sizeof(float),
                            hipMemcpyDeviceToHost),
                  "hipMemcpy");
        for (size_t k = 0; k < n; ++k) out[k] = f[k];
    }
    return out;
}

bool same_scheme(const psst_settings &a, const psst_settings &b) {
    bool eq = a.flux == b.flux && a.order == b.order && a.grad == b.grad && a.limiter == b.limiter &&
              a.time_mode == b.time_mode && a.rk_levels == b.rk_levels && a.cfl == b.cfl && a.gamma == b.gamma &&
              a.rm == b.rm && a.cp == b.cp && a.cv == b.cv && a.precision == b.precision &&
              a.dt_ordering == b.dt_ordering && a.venkat_k == b.venkat_k && a.tile_cells == b.tile_cells &&
              a.tile_threads == b.tile_threads;
    for (int k = 0; k < a.rk_levels && k < 8; ++k) eq = eq && a.rk_coeff[k] == b.rk_coeff[k];
    if (a.time_mode == PSST_DT_UNIFORM) eq = eq && a.uniform_dt == b.uniform_dt;
    return eq;
}

}  // namespace

struct psst_batch {
    psst_ctx *ctx = nullptr;
    std::string err;
    ~psst_batch() {
        if (ctx) psst_ctx_destroy(ctx);
    }
};

extern "C" {

psst_status psst_set_precision(psst_ctx *c, int32_t precision) {
    if (!c) return PSST_BAD_ARG;
    if (precision != 32 && precision != 64) {
        c->err = "precision must be 32 or 64";
        return PSST_BAD_ARG;
    }
    if (precision == c->s.precision) return PSST_OK;
    try {
        const size_t nsb = precision == 64 ? sizeof(double) : sizeof(float);
        if (c->tiled && stage_shm(c, nsb) > 160 * 1024) {
            c->err = "the tile plan does not fit in LDS at this precision";
            return PSST_UNSUPPORTED;
        }
        const size_t nq = (size_t)(c->dim + 3) * c->NT, nw = (size_t)(c->dim + 2) * c->N;
        void *tmp = nullptr;
        hip_check(hipMalloc(&tmp, nq * sizeof(double)), "hipMalloc");
        try {
            void *bufs[3] = {c->q[0], c->q[1], c->W0};
            const size_t cnt[3] = {nq, nq, nw};
            for (int k = 0; k < 3; ++k) {
                if (precision == 64) {
                    hipLaunchKernelGGL((k_convert<float, double>), dim3(blocks((long)cnt[k])), dim3(kBlock), 0,
                                       c->stream, (const float *)bufs[k], (double *)tmp, cnt[k]);
                    hip_check(hipGetLastError(), "convert");
                    hip_check(hipMemcpyAsync(bufs[k], tmp, cnt[k] * sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                              "D2D");
                } else {
                    hipLaunchKernelGGL((k_convert<double, float>), dim3(blocks((long)cnt[k])), dim3(kBlock), 0,
                                       c->stream, (const double *)bufs[k], (float *)tmp, cnt[k]);
                    hip_check(hipGetLastError(), "convert");
                    hip_check(hipMemcpyAsync(bufs[k], tmp, cnt[k] * sizeof(float), hipMemcpyDeviceToDevice, c->stream),
                              "D2D");
                }
            }
            hip_check(hipStreamSynchronize(c->stream), "convert");
        } catch (...) {
            (void)hipFree(tmp);
            throw;
        }
        hip_check(hipFree(tmp), "hipFree");
        c->s.precision = precision;
        c->shm_stage = c->tiled ? stage_shm(c, nsb) : 0;
        c->shm_grad = c->tiled ? grad_shm(c, nsb) : 0;
        destroy_graphs(c);
        return PSST_OK;
    } catch (const std::exception &e) {
        return set_error(c, e, PSST_HIP_ERROR);
    }
}

const char *psst_batch_last_error(const psst_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

psst_status psst_batch_create(int32_t n, const psst_mesh_desc *meshes, const psst_settings *settings,
                              psst_batch **out) {
    if (n <= 0 || !meshes || !settings || !out) {
        g_create_error = "psst_batch_create: null argument or empty batch";
        return PSST_BAD_ARG;
    }
    *out = nullptr;
    try {
        const int dim = meshes[0].dim;
        const psst_settings &s0 = settings[0];
        if (s0.time_mode == PSST_DT_LOCAL1 && s0.dt_ordering == 0)
            throw std::invalid_argument("batches need an order-free time step (time_mode 0 or 2, or dt_ordering 1)");
        int64_t N = 0, G = 0, F = 0, nv = 0, nl = 0, nbc = 0;
        bool lsq = s0.order == 2 && s0.grad == PSST_GRAD_LEAST_SQUARES;
        bool any_shift = false;
        for (int i = 0; i < n; ++i) any_shift = any_shift || meshes[i].face_shift;
        for (int i = 0; i < n; ++i) {
            const psst_mesh_desc &m = meshes[i];
            if (m.dim != dim) throw std::invalid_argument("instances of a batch must share the dimension");
            if (!same_scheme(settings[i], s0)) throw std::invalid_argument("instances of a batch must share the scheme");
            if (m.n_cells <= 0 || m.n_faces <= 0 || m.n_ghosts < 0) throw std::invalid_argument("empty instance mesh");
            if (settings[i].n_bc <= 0 || !settings[i].bc) throw std::invalid_argument("boundary table missing");
            N += m.n_cells;
            G += m.n_ghosts;
            F += m.n_faces;
            nv += m.visit_offsets[m.n_cells];
            nl += m.list_offsets[m.n_cells];
            nbc += settings[i].n_bc;
        }
        const int64_t NT = N + G;
        if (NT >= INT32_MAX / 2 || F >= INT32_MAX / 2 || nv >= INT32_MAX / 2 || nl >= INT32_MAX / 2)
            throw std::invalid_argument("batch too large for 32-bit indices");
        std::vector<double> vol((size_t)NT), cen((size_t)dim * NT), nrm((size_t)dim * F), area((size_t)F),
            beta((size_t)F), lsq_r(lsq ? 6 * (size_t)N : 0), lsq_b(lsq ? (size_t)N : 0),
            fshift(any_shift ? 3 * (size_t)F : 0, 0.0);
        std::vector<int32_t> fc(2 * (size_t)F), voff((size_t)N + 1), vf((size_t)nv), loff((size_t)N + 1), lf((size_t)nl),
            gface((size_t)G), gbc((size_t)G), seg((size_t)n);
        std::vector<psst_bc> bcs;
        bcs.reserve((size_t)nbc);
        auto bs = std::make_unique<BatchState>();
        bs->n = n;
        bs->ws = dim + 2;
        bs->N_tot = (int32_t)N;
        bs->cell_begin.assign(1, 0);
        bs->ghost_begin.assign(1, 0);
        int64_t cb = 0, gb = 0, fb = 0, vb = 0, lb = 0, bb = 0;
        voff[0] = loff[0] = 0;
        for (int i = 0; i < n; ++i) {
            const psst_mesh_desc &m = meshes[i];
            const int64_t Ni = m.n_cells, Gi = m.n_ghosts, Fi = m.n_faces, NTi = Ni + Gi;
            auto uc = [&](int32_t x) -> int32_t {  // instance cell -> union canonical cell
                return (int32_t)(x < Ni ? cb + x : N + gb + (x - Ni));
            };
            seg[(size_t)i] = (int32_t)cb;
            for (int64_t x = 0; x < NTi; ++x) {
                const int32_t u = uc((int32_t)x);
                vol[(size_t)u] = m.cell_volume[x];
                for (int k = 0; k < dim; ++k) cen[(size_t)k * NT + u] = m.cell_centre[(size_t)k * NTi + x];
            }
            for (int64_t f = 0; f < Fi; ++f) {
                for (int k = 0; k < dim; ++k) nrm[(size_t)k * F + fb + f] = m.face_normal[(size_t)k * Fi + f];
                if (m.face_shift)
                    for (int k = 0; k < 3; ++k) fshift[(size_t)k * F + fb + f] = m.face_shift[(size_t)k * Fi + f];
                area[(size_t)(fb + f)] = m.face_area[f];
                beta[(size_t)(fb + f)] = m.face_beta[f];
                const int32_t a = m.face_cells[f], b = m.face_cells[Fi + f];
                if (a < 0 || a >= NTi || b < 0 || b >= NTi) throw std::invalid_argument("face_cells out of range");
                fc[(size_t)(fb + f)] = uc(a);
                fc[(size_t)(F + fb + f)] = uc(b);
            }
            for (int64_t x = 0; x < Ni; ++x) {
                for (int32_t k = m.visit_offsets[x]; k < m.visit_offsets[x + 1]; ++k)
                    vf[(size_t)(vb + k)] = (int32_t)(fb + m.visit_faces[k]);
                for (int32_t k = m.list_offsets[x]; k < m.list_offsets[x + 1]; ++k)
                    lf[(size_t)(lb + k)] = (int32_t)(fb + m.list_faces[k]);
                voff[(size_t)(cb + x + 1)] = (int32_t)(vb + m.visit_offsets[x + 1]);
                loff[(size_t)(cb + x + 1)] = (int32_t)(lb + m.list_offsets[x + 1]);
                if (lsq) {
                    if (!m.lsq_r || !m.lsq_beta) throw std::invalid_argument("least squares needs lsq_r and lsq_beta");
                    for (int r = 0; r < 6; ++r) lsq_r[(size_t)r * N + cb + x] = m.lsq_r[(size_t)r * Ni + x];
                    lsq_b[(size_t)(cb + x)] = m.lsq_beta[x];
                }
            }
            for (int64_t g = 0; g < Gi; ++g) {
                gface[(size_t)(gb + g)] = (int32_t)(fb + m.ghost_face[g]);
                if (m.ghost_bc[g] < 0 || m.ghost_bc[g] >= settings[i].n_bc) throw std::invalid_argument("ghost_bc out of range");
                gbc[(size_t)(gb + g)] = (int32_t)(bb + m.ghost_bc[g]);
            }
            for (int k = 0; k < settings[i].n_bc; ++k) bcs.push_back(settings[i].bc[k]);
            cb += Ni;
            gb += Gi;
            fb += Fi;
            vb += m.visit_offsets[Ni];
            lb += m.list_offsets[Ni];
            bb += settings[i].n_bc;
            bs->cell_begin.push_back((int32_t)cb);
            bs->ghost_begin.push_back((int32_t)gb);
        }
        double vmin = meshes[0].min_volume;
        for (int i = 1; i < n; ++i) vmin = std::min(vmin, meshes[i].min_volume);
        psst_mesh_desc um{};
        um.dim = dim;
        um.n_cells = (int32_t)N;
        um.n_ghosts = (int32_t)G;
        um.n_faces = (int32_t)F;
        um.cell_volume = vol.data();
        um.cell_centre = cen.data();
        um.face_normal = nrm.data();
        um.face_area = area.data();
        um.face_cells = fc.data();
        um.face_beta = beta.data();
        um.visit_offsets = voff.data();
        um.visit_faces = vf.data();
        um.list_offsets = loff.data();
        um.list_faces = lf.data();
        um.ghost_face = gface.data();
        um.ghost_bc = gbc.data();
        um.lsq_r = lsq ? lsq_r.data() : nullptr;
        um.lsq_beta = lsq ? lsq_b.data() : nullptr;
        um.min_volume = vmin;
        um.face_shift = any_shift ? fshift.data() : nullptr;
        psst_settings us = s0;
        us.fused = 1;
        us.bc = bcs.data();
        us.n_bc = (int32_t)bcs.size();
        psst_ctx *c = nullptr;
        psst_status st = create_impl(&um, &us, seg, &c);
        if (st != PSST_OK) return st;
        auto bt = std::make_unique<psst_batch>();
        bt->ctx = c;
        if (!c->fused) {
            psst_ctx_destroy(c);
            bt->ctx = nullptr;
            throw std::invalid_argument("the batch's tiles do not fit in LDS (use smaller tile_cells)");
        }
        // tiles never straddle instances: the tiles of instance i are those
        // whose first cell lies in [cell_begin[i], cell_begin[i+1])
        bs->tile_begin.assign(1, 0);
        std::vector<int32_t> tile_inst((size_t)c->n_tiles);
        for (int t = 0, i = 0; t < c->n_tiles; ++t) {
            const int32_t cb0 = c->tile_cell_begin[(size_t)t];
            while (cb0 >= bs->cell_begin[(size_t)i + 1]) {
                bs->tile_begin.push_back(t);
                ++i;
            }
            if (c->tile_cell_begin[(size_t)t + 1] > bs->cell_begin[(size_t)i + 1])
                throw std::logic_error("a tile straddles two instances");
            tile_inst[(size_t)t] = i;
        }
        while ((int)bs->tile_begin.size() < n + 1) bs->tile_begin.push_back(c->n_tiles);
        bs->c2d.resize((size_t)NT);
        for (int64_t d = 0; d < NT; ++d) bs->c2d[(size_t)c->d2c[(size_t)d]] = (int32_t)d;
        bs->have_state.assign((size_t)n, 0);
        bs->enabled.assign((size_t)n, 1);
        bs->w_dirty.assign((size_t)n, 0);
        bs->buf_of.assign((size_t)n, 0);
        const int ws = dim + 2;
        psst_ctx *cx = c;
        bs->tile_inst = dev_upload(cx, tile_inst.data(), tile_inst.size());
        bs->tile_begin_dev = dev_upload(cx, bs->tile_begin.data(), bs->tile_begin.size());
        bs->list = dev_alloc<int>(cx, (size_t)c->n_tiles);
        std::vector<int> zeros_i((size_t)std::max(n, 8 * 9), 0), minus1((size_t)n, -1);
        bs->count = dev_upload(cx, zeros_i.data(), 1);
        bs->iter = dev_upload(cx, zeros_i.data(), 1);
        bs->active = dev_upload(cx, zeros_i.data(), (size_t)n);
        bs->frozen_buf = dev_upload(cx, zeros_i.data(), (size_t)n);
        bs->conv = dev_upload(cx, minus1.data(), (size_t)n);
        std::vector<unsigned long long> zres((size_t)n * ws * kResSlots, 0ull);
        bs->res = dev_upload(cx, zres.data(), zres.size());
        std::vector<double> mres((size_t)n * ws, -1.0), zd((size_t)n, 0.0);
        bs->res0 = dev_upload(cx, mres.data(), mres.size());
        bs->drop = dev_upload(cx, zd.data(), zd.size());
        bs->hist = dev_alloc<double>(cx, (size_t)bs->H * n * ws);
        c->batch = std::move(bs);
        // after the move, so that the context's destroy frees them on any
        // later failure
        BatchState &b = *c->batch;
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&b.h_flags), sizeof(int) * (3 * (size_t)n + 1)),
                  "hipHostMalloc");
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&b.h_ring), sizeof(double) * (size_t)b.H * n * ws),
                  "hipHostMalloc");
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&b.h_drop), sizeof(double) * (size_t)n), "hipHostMalloc");
        *out = bt.release();
        return PSST_OK;
    } catch (const std::exception &e) {
        g_create_error = e.what();
        return dynamic_cast<const HipError *>(&e) ? PSST_HIP_ERROR : PSST_BAD_ARG;
    }
}

psst_ctx *psst_batch_context(psst_batch *b) { return b ? b->ctx : nullptr; }

psst_status psst_batch_set_state(psst_batch *bt, int32_t inst, const double *prim) {
    if (!bt || !prim) return PSST_BAD_ARG;
    psst_ctx *c = bt->ctx;
    BatchState &b = *c->batch;
    if (inst < 0 || inst >= b.n) {
        bt->err = "instance index out of range";
        return PSST_BAD_ARG;
    }
    try {
        const int32_t cb = b.cell_begin[(size_t)inst], Ni = b.cell_begin[(size_t)inst + 1] - cb;
        const int32_t gb = b.ghost_begin[(size_t)inst], Gi = b.ghost_begin[(size_t)inst + 1] - gb;
        const size_t NTi = (size_t)Ni + Gi;
        std::vector<double> vi((size_t)Ni), vg((size_t)Gi);
        for (int v = 0; v < c->dim + 3; ++v) {
            for (int32_t x = 0; x < Ni; ++x) vi[(size_t)(b.c2d[(size_t)(cb + x)] - cb)] = prim[(size_t)v * NTi + x];
            for (int32_t g = 0; g < Gi; ++g) vg[(size_t)g] = prim[(size_t)v * NTi + Ni + g];
            for (int k = 0; k < 2; ++k) {
                upload_values(c, c->q[k], (size_t)v * c->NT + cb, vi);
                upload_values(c, c->q[k], (size_t)v * c->NT + b.N_tot + gb, vg);
            }
        }
        const int m1 = -1;
        hip_check(hipMemcpy(b.conv + inst, &m1, sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
        b.have_state[(size_t)inst] = 1;
        b.buf_of[(size_t)inst] = c->cur;
        b.w_dirty[(size_t)inst] = 1;
        c->have_state = true;
        return PSST_OK;
    } catch (const std::exception &e) {
        bt->err = e.what();
        return PSST_HIP_ERROR;
    }
}

psst_status psst_batch_get_state(psst_batch *bt, int32_t inst, double *prim) {
    if (!bt || !prim) return PSST_BAD_ARG;
    psst_ctx *c = bt->ctx;
    BatchState &b = *c->batch;
    if (inst < 0 || inst >= b.n) {
        bt->err = "instance index out of range";
        return PSST_BAD_ARG;
    }
    try {
        const int32_t cb = b.cell_begin[(size_t)inst], Ni = b.cell_begin[(size_t)inst + 1] - cb;
        const int32_t gb = b.ghost_begin[(size_t)inst], Gi = b.ghost_begin[(size_t)inst + 1] - gb;
        const size_t NTi = (size_t)Ni + Gi;
        const void *src = c->q[b.buf_of[(size_t)inst]];
        for (int v = 0; v < c->dim + 3; ++v) {
            std::vector<double> vi = download_values(c, src, (size_t)v * c->NT + cb, (size_t)Ni);
            std::vector<double> vg = download_values(c, src, (size_t)v * c->NT + b.N_tot + gb, (size_t)Gi);
            for (int32_t x = 0; x < Ni; ++x) prim[(size_t)v * NTi + x] = vi[(size_t)(b.c2d[(size_t)(cb + x)] - cb)];
            for (int32_t g = 0; g < Gi; ++g) prim[(size_t)v * NTi + Ni + g] = vg[(size_t)g];
        }
        return PSST_OK;
    } catch (const std::exception &e) {
        bt->err = e.what();
        return PSST_HIP_ERROR;
    }
}

psst_status psst_batch_set_residual_reference(psst_batch *bt, int32_t inst, const double *res0) {
    if (!bt || !res0) return PSST_BAD_ARG;
    psst_ctx *c = bt->ctx;
    BatchState &b = *c->batch;
    if (inst < 0 || inst >= b.n) {
        bt->err = "instance index out of range";
        return PSST_BAD_ARG;
    }
    if (hipMemcpy(b.res0 + (size_t)inst * b.ws, res0, sizeof(double) * b.ws, hipMemcpyHostToDevice) != hipSuccess) {
        bt->err = "hipMemcpy failed";
        return PSST_HIP_ERROR;
    }
    return PSST_OK;
}

psst_status psst_batch_set_enabled(psst_batch *bt, int32_t inst, int32_t enabled) {
    if (!bt) return PSST_BAD_ARG;
    BatchState &b = *bt->ctx->batch;
    if (inst < 0 || inst >= b.n) {
        bt->err = "instance index out of range";
        return PSST_BAD_ARG;
    }
    b.enabled[(size_t)inst] = enabled != 0;
    return PSST_OK;
}

psst_status psst_batch_set_graph(psst_batch *bt, int32_t enable) {
    if (!bt) return PSST_BAD_ARG;
    bt->ctx->batch->use_graph = enable != 0;
    return PSST_OK;
}

psst_status psst_batch_run(psst_batch *bt, int32_t max_iter, double tol, int32_t *iters, int32_t *status,
                           double *drop_out, double *res_hist) {
    if (!bt) return PSST_BAD_ARG;
    psst_ctx *c = bt->ctx;
    BatchState &b = *c->batch;
    for (int i = 0; i < b.n; ++i)
        if (!b.have_state[(size_t)i]) {
            bt->err = "psst_batch_run: instance " + std::to_string(i) + " has no state";
            return PSST_BAD_ARG;
        }
    if (max_iter <= 0) return PSST_OK;
    const int n = b.n, ws = b.ws;
    try {
        batch_gather_states(c);
        // conservative state of the instances given a new state (only theirs:
        // the others keep the exact w of their last iteration)
        for (int i = 0; i < n; ++i) {
            if (!b.w_dirty[(size_t)i]) continue;
            const int c0 = b.cell_begin[(size_t)i], c1 = b.cell_begin[(size_t)i + 1];
            if (c->dim == 2) c->s.precision == 64 ? run_set_w<2, double>(c, c0, c1) : run_set_w<2, float>(c, c0, c1);
            else c->s.precision == 64 ? run_set_w<3, double>(c, c0, c1) : run_set_w<3, float>(c, c0, c1);
            b.w_dirty[(size_t)i] = 0;
        }
        // Every transfer of the run goes through the pinned host buffers
        // allocated with the batch (h_flags: active [n], active-tile count,
        // conv [n], frozen_buf [n]; h_drop; h_ring), in stream order.
        int *h_active = b.h_flags, *h_count = b.h_flags + n, *h_conv = b.h_flags + n + 1,
            *h_frozen = b.h_flags + 2 * n + 1;
        hip_check(hipMemcpyAsync(h_conv, b.conv, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream), "hipMemcpy");
        hip_check(hipStreamSynchronize(c->stream), "psst_batch_run");
        // every instance's latest state is in the current buffer now; one
        // that does not run keeps it there
        const int cur_start = c->cur;
        std::vector<char> active_start((size_t)n);
        for (int i = 0; i < n; ++i) {
            h_active[i] = h_conv[i] != -2 && b.enabled[(size_t)i];
            active_start[(size_t)i] = (char)h_active[i];
            if (h_conv[i] != -2) h_conv[i] = -1;
        }
        hip_check(hipMemcpyAsync(b.active, h_active, sizeof(int) * n, hipMemcpyHostToDevice, c->stream), "hipMemcpy");
        hip_check(hipMemcpyAsync(b.conv, h_conv, sizeof(int) * n, hipMemcpyHostToDevice, c->stream), "hipMemcpy");
        launch_check(c, tol, c->cur, 1);
        hip_check(hipMemcpyAsync(h_count, b.count, sizeof(int), hipMemcpyDeviceToHost, c->stream), "hipMemcpy");
        hip_check(hipStreamSynchronize(c->stream), "psst_batch_run");
        int count = *h_count;
        // the active list only shrinks during a run, so a capacity chosen at
        // a chunk's start holds it for the whole chunk
        b.grid_tiles = batch_grid_tiles(count, c->n_tiles);
        c->bd = BatchDev{b.list, b.count, b.tile_inst, b.res};
        const int64_t it0 = b.iter_host;
        int done = 0;
        bool any = true, first = true;
        std::vector<char> active_now(active_start);
        while (done < max_iter && any) {
            const int m = std::min(b.H, max_iter - done);
            for (int k = 0; k < m; ++k) {
                if (first || !b.use_graph || c->timing) batch_iteration(c, tol);
                else batch_graph_iteration(c, tol);
                first = false;
            }
            if (res_hist)
                hip_check(hipMemcpyAsync(b.h_ring, b.hist, sizeof(double) * (size_t)b.H * n * ws,
                                         hipMemcpyDeviceToHost, c->stream),
                          "hipMemcpy");
            hip_check(hipMemcpyAsync(h_active, b.active, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream),
                      "hipMemcpy");
            hip_check(hipMemcpyAsync(h_count, b.count, sizeof(int), hipMemcpyDeviceToHost, c->stream), "hipMemcpy");
            hip_check(hipStreamSynchronize(c->stream), "psst_batch_run");
            for (int i = 0; i < n; ++i) active_now[(size_t)i] = (char)(h_active[i] != 0);
            count = *h_count;
            b.grid_tiles = batch_grid_tiles(count, c->n_tiles);
            if (c->timing) flush_timing(c);
            if (res_hist)
                for (int k = 0; k < m; ++k) {
                    const size_t row = (size_t)((b.iter_host + k) % b.H);
                    std::memcpy(res_hist + (size_t)(done + k) * n * ws, b.h_ring + row * n * ws,
                                sizeof(double) * n * ws);
                }
            b.iter_host += m;
            done += m;
            any = false;
            for (int i = 0; i < n; ++i) any = any || active_now[(size_t)i];
        }
        c->bd = BatchDev{};
        b.grid_tiles = 0;
        hip_check(hipMemcpyAsync(h_conv, b.conv, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream), "hipMemcpy");
        hip_check(hipMemcpyAsync(h_frozen, b.frozen_buf, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream),
                  "hipMemcpy");
        hip_check(hipMemcpyAsync(b.h_drop, b.drop, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream), "hipMemcpy");
        hip_check(hipStreamSynchronize(c->stream), "psst_batch_run");
        psst_status worst = PSST_OK;
        for (int i = 0; i < n; ++i) {
            const int cv = h_conv[i];
            // the buffer that holds the instance's latest state: the current
            // one if it still runs, the one of its stopping iteration if it
            // stopped in this run, the run's starting buffer if it never ran
            b.buf_of[(size_t)i] = active_now[(size_t)i] ? c->cur : (active_start[(size_t)i] ? h_frozen[i] : cur_start);
            psst_status si = PSST_OK;
            if (cv == -2) si = PSST_NAN_DETECTED;
            else if (cv < 0 && tol > 0.0) si = PSST_NO_CONVERGENCE;
            if (!b.enabled[(size_t)i] && cv != -2) si = PSST_OK;
            if (iters) iters[i] = cv >= 0 ? (int32_t)(cv - it0 + 1) : (b.enabled[(size_t)i] ? done : 0);
            if (status) status[i] = si;
            if (drop_out) drop_out[i] = b.h_drop[i];
            if (si == PSST_NAN_DETECTED) worst = PSST_NAN_DETECTED;
            else if (si == PSST_NO_CONVERGENCE && worst == PSST_OK) worst = PSST_NO_CONVERGENCE;
        }
        if (worst == PSST_NAN_DETECTED) bt->err = "non-finite residual in at least one instance";
        else if (worst == PSST_NO_CONVERGENCE) bt->err = "at least one instance did not reach the tolerance";
        return worst;
    } catch (const std::exception &e) {
        c->bd = BatchDev{};
        b.grid_tiles = 0;
        bt->err = e.what();
        return PSST_HIP_ERROR;
    }
}

psst_status psst_batch_set_precision(psst_batch *bt, int32_t precision) {
    if (!bt) return PSST_BAD_ARG;
    psst_status st = psst_set_precision(bt->ctx, precision);
    if (st != PSST_OK) bt->err = bt->ctx->err;
    return st;
}

psst_status psst_batch_info(const psst_batch *bt, int64_t *info) {
    if (!bt || !info) return PSST_BAD_ARG;
    const psst_ctx *c = bt->ctx;
    info[0] = c->batch->n;
    info[1] = c->N;
    info[2] = c->G;
    info[3] = c->F;
    info[4] = c->n_tiles;
    info[5] = c->tile_cells;
    info[6] = c->tile_threads;
    info[7] = c->batch->iter_host;
    return PSST_OK;
}

void psst_batch_destroy(psst_batch *bt) { delete bt; }

psst_status psst_autotune(const psst_mesh_desc *m, const psst_settings *s, const double *prim, int32_t trial_iters,
                          int32_t max_cand, int32_t *cand, double *ms_per_iter, int32_t *n_cand, int32_t *best) {
    if (!m || !s || !prim || trial_iters <= 0 || !best) {
        g_create_error = "psst_autotune: bad argument";
        return PSST_BAD_ARG;
    }
    // tile shapes: cells per tile x threads per workgroup (the kernels are
    // built for 256 and 512 threads)
    static const int shapes[][2] = {{128, 256}, {160, 256}, {192, 256}, {224, 256}, {256, 256}, {256, 512},
                                    {384, 256}, {384, 512}, {512, 256}, {512, 512}, {768, 512}, {1024, 512}};
    double best_ms = 1e300;
    int k_out = 0;
    best[0] = best[1] = 0;
    for (const auto &sh : shapes) {
        psst_settings t = *s;
        t.fused = 1;
        t.tile_cells = sh[0];
        t.tile_threads = sh[1];
        psst_ctx *c = nullptr;
        double ms = -1.0;
        if (create_impl(m, &t, {}, &c) == PSST_OK) {
            if (c->fused && psst_set_state(c, prim) == PSST_OK) {
                int32_t done = 0;
                if (psst_run(c, 2, 0.0, nullptr, &done) == PSST_OK) {
                    const auto t0 = std::chrono::steady_clock::now();
                    if (psst_run(c, trial_iters, 0.0, nullptr, &done) == PSST_OK)
                        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
                             trial_iters;
                }
            }
            psst_ctx_destroy(c);
        }
        if (cand && k_out < max_cand) {
            cand[2 * k_out] = sh[0];
            cand[2 * k_out + 1] = sh[1];
        }
        if (ms_per_iter && k_out < max_cand) ms_per_iter[k_out] = ms;
        ++k_out;
        if (ms > 0.0 && ms < best_ms) {
            best_ms = ms;
            best[0] = sh[0];
            best[1] = sh[1];
        }
    }
    if (n_cand) *n_cand = std::min(k_out, std::max(max_cand, 0));
    if (best[0] == 0) {
        g_create_error = "psst_autotune: no tile shape ran";
        return PSST_UNSUPPORTED;
    }
    return PSST_OK;
}

}  // extern "C"
