// Two-site DMRG sweep driver on one GPU.
//
// Restates dmrg_two_site (reference src/algorithms/dmrg.mojo:377-483),
// DMRGWorkspace (:39-90) and dmrg_local_update_two_site (:759-844) with every
// tensor resident in HBM: MPS sites, all N+1 left and right environments, the
// Krylov basis and the SVD factors.  Per local update the host reads back only
// the Lanczos scalars and the singular values (for the truncation rule); the
// reference instead moves every Krylov vector and the full SVD operands over
// PCIe (dmrg.mojo:640-689, dense_tensor.mojo:1514-1590).
//
// Storage is sized once for the largest bond each position can reach,
// min(d^l, d^(N-l), chi_max) (or the initial bond when larger), so a sweep
// performs no allocation.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>

#include "ct_internal.h"

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// The Lanczos solve of one bond as a HIP graph.  key: everything the
// recorded kernel arguments depend on (shapes, precision, solver settings and
// the bond's device buffers); gen: the handle's workspace generation at
// capture (a reallocated workspace makes the recorded addresses stale).
struct BondGraph {
    std::array<int64_t, 16> seen_key{};  // key of the last eager run: every workspace is large enough for it
    bool seen = false;
    bool no_capture = false;             // capture failed once: stay eager for this bond
    std::array<int64_t, 16> exec_key{};  // key the graph was captured for
    uint64_t gen = 0;
    hipGraphExec_t exec = nullptr;
};

struct ct_dmrg_s {
    ct::Handle* h = nullptr;
    ct_dmrg_params p{};
    int N = 0;
    int64_t d = 0;
    std::vector<int64_t> w, D, Dcap;
    std::vector<double*> W, W2, A, Lenv, Renv;
    double* theta = nullptr;
    double* S = nullptr;
    ct_dmrg_stats st{};
    std::vector<double> sweep_E;
    std::vector<BondGraph> graphs;
    hipStream_t gstream = nullptr;  // capture stream (the caller's stream may be the null stream)

    ~ct_dmrg_s() {
        if (h) (void)hipStreamSynchronize(h->stream);
        if (gstream) (void)hipStreamSynchronize(gstream);
        for (BondGraph& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (gstream) (void)hipStreamDestroy(gstream);
        for (auto* v : {&W, &W2, &A, &Lenv, &Renv})
            for (double* q : *v)
                if (q) (void)hipFree(q);
        if (theta) (void)hipFree(theta);
        if (S) (void)hipFree(S);
    }
};

namespace {

double matvec_flops(int64_t Dl, int64_t d, int64_t Dr, int64_t wL, int64_t wM, int64_t wR) {
    // reference contraction order (dmrg.mojo:310-372)
    const double a = 2.0 * wL * Dl * Dl * d * d * Dr;
    const double b = 2.0 * Dl * d * Dr * wL * d * d * wM;
    const double c = 2.0 * Dl * Dr * d * wM * d * d * wR;
    const double e = 2.0 * Dl * d * d * wR * Dr * Dr;
    return a + b + c + e;
}

double env_flops(int64_t Dl, int64_t d, int64_t Dr, int64_t wL, int64_t wR) {
    return 2.0 * wL * Dl * Dl * d * Dr + 2.0 * Dl * Dr * wL * d * d * wR + 2.0 * Dr * wR * Dl * d * Dr;
}

// FP64 copy of a caller's FP64 or float32 device buffer
int copy_in(ct::Handle* h, int dtype, const void* src, double* dst, int64_t n) {
    if (dtype == CT_F64) return ct::hip_rc(hipMemcpyAsync(dst, src, (size_t)n * 8, hipMemcpyDeviceToDevice, h->stream));
    return ct::convert(h, CT_F32, src, CT_F64, dst, n);
}

int alloc_d(double** p, int64_t n) {
    if (n < 1) n = 1;
    CT_HIP(hipMalloc(p, (size_t)n * sizeof(double)));
    return 0;
}

int build_envs(ct_dmrg_s* s) {
    ct::Handle* h = s->h;
    const int N = s->N;
    // R[N] = identity [w_N, D_N, D_N] (environments.mojo:441-461)
    CT_CHECK(ct::set_identity_env(h, CT_F64, s->Renv[N], s->w[N], s->D[N]));
    for (int i = N - 1; i >= 0; --i)
        CT_CHECK(ct::env_right(h, s->p.env_prec, CT_F64, s->Renv[i + 1], s->A[i], s->W[i], s->Renv[i], s->D[i], s->d,
                               s->D[i + 1], s->w[i], s->w[i + 1]));
    // L[0] = identity [w_0, D_0, D_0] (dmrg.mojo:70-90)
    CT_CHECK(ct::set_identity_env(h, CT_F64, s->Lenv[0], s->w[0], s->D[0]));
    return 0;
}

// H_eff preparation (both precisions) and the Lanczos solve of bond i,
// enqueued on h->stream without any host synchronisation.
int enqueue_local_solve(ct_dmrg_s* s, int i, const ct::HeffDims& hd) {
    ct::Handle* h = s->h;
    ct::HeffEnv env, env64;
    CT_CHECK(ct::heff_prepare(h, s->p.prec, CT_F64, s->Lenv[i], s->Renv[i + 2], s->W2[i], hd, &env));
    CT_CHECK(ct::heff_prepare(h, CT_PREC_FP64, CT_F64, s->Lenv[i], s->Renv[i + 2], s->W2[i], hd, &env64));
    return ct::lanczos_enqueue(h, env, &env64, s->theta, s->p.max_krylov_iter, s->p.krylov_tol, s->p.jacobi_max_rot,
                               s->p.refine);
}

// The local solve, launched kernel by kernel the first time a bond's key is
// seen, captured into a graph the second time and replayed afterwards (at a
// fixed bond dimension every sweep repeats every key).  A capture that fails
// for any reason falls back to launching the kernels.
int local_solve(ct_dmrg_s* s, int i, const ct::HeffDims& hd, ct::LanczosResult* lr) {
    ct::Handle* h = s->h;
    if (!s->p.graphs) {
        CT_CHECK(enqueue_local_solve(s, i, hd));
        return ct::lanczos_finish(h, s->p.refine, true, lr);
    }
    std::array<int64_t, 16> key{};
    double tol = s->p.krylov_tol;
    int64_t tol_bits;
    std::memcpy(&tol_bits, &tol, sizeof tol_bits);
    key = {hd.Dl, hd.Dr, hd.wL, hd.wM, hd.wR, s->d, s->p.prec, s->p.refine, s->p.max_krylov_iter,
           s->p.jacobi_max_rot, tol_bits, (int64_t)(uintptr_t)s->theta, (int64_t)(uintptr_t)s->Lenv[i],
           (int64_t)(uintptr_t)s->Renv[i + 2], (int64_t)(uintptr_t)s->W2[i], 0};
    BondGraph& g = s->graphs[i];
    const bool replay = g.exec && g.exec_key == key && g.gen == h->ws_gen;
    if (!replay) {
        if (g.no_capture || !g.seen || g.seen_key != key) {
            CT_CHECK(enqueue_local_solve(s, i, hd));
            if (!g.no_capture) {
                g.seen_key = key;
                g.seen = true;
            }
            return ct::lanczos_finish(h, s->p.refine, true, lr);
        }
        // capture (the eager run with this key sized every workspace)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
        hipStream_t user = h->stream;
        hipGraph_t graph = nullptr;
        int rc = ct::hip_rc(hipStreamBeginCapture(s->gstream, hipStreamCaptureModeRelaxed));
        if (rc == 0) {
            h->stream = s->gstream;
            rc = enqueue_local_solve(s, i, hd);
            h->stream = user;
            const hipError_t e = hipStreamEndCapture(s->gstream, &graph);
            if (rc == 0 && e != hipSuccess) rc = ct::hip_rc(e);
        }
        if (rc == 0 && graph) rc = ct::hip_rc(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
        if (graph) (void)hipGraphDestroy(graph);
        if (rc != 0 || !g.exec) {
            if (s->p.verbose)
                fprintf(stderr, "bond %d: graph capture failed (status %d: %s); launching its kernels instead\n", i, rc,
                        rc != 0 ? ct_last_error() : "no graph");
            (void)hipGetLastError();
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
            g.no_capture = true;
            CT_CHECK(enqueue_local_solve(s, i, hd));
            return ct::lanczos_finish(h, s->p.refine, true, lr);
        }
        g.exec_key = key;
        g.gen = h->ws_gen;
    }
    // replay on the caller's stream: ordered after theta, before the truncation
    CT_HIP(hipGraphLaunch(g.exec, h->stream));
    s->st.graph_replays += 1;
    return ct::lanczos_finish(h, s->p.refine, true, lr);
}

int local_update(ct_dmrg_s* s, int i, bool l2r, double* E_out) {
    ct::Handle* h = s->h;
    const int64_t d = s->d;
    const int64_t Dl = s->D[i], Dm = s->D[i + 1], Dr = s->D[i + 2];
    const int64_t wL = s->w[i], wM = s->w[i + 1], wR = s->w[i + 2];
    const int64_t m = Dl * d, n = d * Dr;
    double t0 = now_s();
    {   // build_two_site_theta (dmrg.mojo:93-140)
        ct::GemmDesc g;
        g.M = (int)m; g.N = (int)n; g.K = (int)Dm;
        g.a_kc = 1; g.sAm = Dm; g.sAk = 1;
        g.b_kc = 0; g.sBk = n; g.sBn = 1;
        g.ldc = n;
        CT_CHECK(ct::gemm(h, CT_PREC_FP64, CT_F64, s->A[i], CT_F64, s->A[i + 1], CT_F64, s->theta, g));
    }
    double t1 = now_s();
    ct::HeffDims hd{Dl, Dl, d, Dr, Dr, wL, wM, wR};
    ct::LanczosResult lr;
    CT_CHECK(local_solve(s, i, hd, &lr));
    double t2 = now_s();
    // split_two_site_theta_svd (dmrg.mojo:143-252): A_i = U (S) , A_{i+1} = (S) Vt
    ct::SvdResult sr;
    CT_CHECK(ct::theta_split(h, s->p.svd_mode, s->theta, m, n, s->p.chi_max, s->p.eps_trunc, l2r, s->A[i],
                             s->A[i + 1], &sr));
    if (sr.chi <= 0) {
        ct::set_last_error("truncation kept no singular value");
        return CT_STATUS_SVD_FAILED;
    }
    const int64_t chi = sr.chi;
    s->D[i + 1] = chi;
    CT_HIP(hipStreamSynchronize(h->stream));
    double t3 = now_s();
    if (l2r)
        CT_CHECK(ct::env_left(h, s->p.env_prec, CT_F64, s->Lenv[i], s->A[i], s->W[i], s->Lenv[i + 1], Dl, d, chi, wL,
                              wM));
    else
        CT_CHECK(ct::env_right(h, s->p.env_prec, CT_F64, s->Renv[i + 2], s->A[i + 1], s->W[i + 1], s->Renv[i + 1],
                               chi, d, Dr, wM, wR));
    CT_HIP(hipStreamSynchronize(h->stream));
    double t4 = now_s();

    ct_dmrg_stats& st = s->st;
    st.t_theta += t1 - t0;
    st.t_lanczos += t2 - t1;
    st.t_svd += t3 - t2;
    st.t_env += t4 - t3;
    st.local_updates += 1;
    const int nmv = lr.matvecs;  // Krylov matvecs + refinement
    st.matvecs += nmv;
    st.kept_start += lr.kept_start;
    st.matvec_flops += nmv * matvec_flops(Dl, d, Dr, wL, wM, wR);
    st.env_flops += l2r ? env_flops(Dl, d, chi, wL, wM) : env_flops(chi, d, Dr, wM, wR);
    st.trunc_weight_sum += sr.discarded;
    if (sr.discarded > st.trunc_weight_max) st.trunc_weight_max = sr.discarded;
    if (chi > st.max_bond) st.max_bond = chi;
    *E_out = lr.E0;
    if (s->p.verbose)
        fprintf(stderr, "  site %d %s E=%.15g (lanczos %.15g, iters %d) chi=%lld trunc=%.3e\n", i, l2r ? "L->R" : "R->L",
                lr.E0, lr.E_lanczos, lr.num_iter, (long long)chi, sr.discarded);
    return 0;
}

}  // namespace

extern "C" {

int ct_dmrg_create(ct_handle_t h, const ct_dmrg_params* params, int nsites, int64_t d, int dtype,
                   const int64_t* mpo_dims, const void* const* mpo_sites, const int64_t* mps_dims,
                   const void* const* mps_sites, ct_dmrg_t* out) {
    CT_ARG(h != nullptr, 1, "handle is NULL");
    CT_ARG(params != nullptr, 2, "params is NULL");
    CT_ARG(nsites >= 2, 3, "DMRG requires at least 2 sites");
    CT_ARG(d >= 1, 4, "d must be >= 1");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 5, "unknown dtype");
    CT_ARG(params->chi_max >= 1, 2, "chi_max must be >= 1");
    CT_ARG(params->max_krylov_iter >= 1 && params->max_krylov_iter <= 64, 2, "max_krylov_iter must be in [1, 64]");
    CT_ARG(params->prec >= 0 && params->prec <= CT_PREC_BF16X6, 2, "unknown precision");
    CT_ARG(params->env_prec >= 0 && params->env_prec <= CT_PREC_BF16X6, 2, "unknown environment precision");
    CT_ARG(params->jacobi_max_rot >= -1, 2, "jacobi_max_rot must be >= -1");
    CT_ARG(params->svd_mode == CT_SVD_GESVD || params->svd_mode == CT_SVD_GRAM, 2, "unknown svd_mode");
    ct_dmrg_s* s = new ct_dmrg_s();
    s->h = h;
    s->p = *params;
    s->N = nsites;
    s->d = d;
    const int N = nsites;
    s->w.assign(mpo_dims, mpo_dims + N + 1);
    s->D.assign(mps_dims, mps_dims + N + 1);
    // Largest bond each position can reach: min(chi_max, d * neighbour) grown
    // from both ends (or the initial bond when larger).
    s->Dcap.resize(N + 1);
    {
        std::vector<int64_t> cl(N + 1), cr(N + 1);
        auto grow = [&](int64_t prev) { return prev > params->chi_max / d ? params->chi_max : std::min(params->chi_max, prev * d); };
        cl[0] = s->D[0];
        for (int l = 1; l <= N; ++l) cl[l] = std::max(s->D[l], grow(cl[l - 1]));
        cr[N] = s->D[N];
        for (int l = N - 1; l >= 0; --l) cr[l] = std::max(s->D[l], grow(cr[l + 1]));
        for (int l = 0; l <= N; ++l) s->Dcap[l] = std::max(s->D[l], std::min(cl[l], cr[l]));
    }
    s->graphs.assign(N, BondGraph());
    s->W.assign(N, nullptr);
    s->W2.assign(N, nullptr);
    s->A.assign(N, nullptr);
    s->Lenv.assign(N + 1, nullptr);
    s->Renv.assign(N + 1, nullptr);
    int rc = 0;
    if (params->graphs) {  // default priority, like every stream of the library
        rc = ct::hip_rc(hipStreamCreateWithFlags(&s->gstream, hipStreamNonBlocking));
    }
    int64_t theta_cap = 1, s_cap = 1;
    for (int i = 0; i < N && rc == 0; ++i) {
        const int64_t nw = s->w[i] * d * d * s->w[i + 1];
        rc = alloc_d(&s->W[i], nw);
        if (rc == 0) rc = copy_in(h, dtype, mpo_sites[i], s->W[i], nw);
        if (rc == 0) rc = alloc_d(&s->A[i], s->Dcap[i] * d * s->Dcap[i + 1]);
        if (rc == 0) rc = copy_in(h, dtype, mps_sites[i], s->A[i], s->D[i] * d * s->D[i + 1]);
    }
    for (int l = 0; l <= N && rc == 0; ++l) {
        rc = alloc_d(&s->Lenv[l], s->w[l] * s->Dcap[l] * s->Dcap[l]);
        if (rc == 0) rc = alloc_d(&s->Renv[l], s->w[l] * s->Dcap[l] * s->Dcap[l]);
    }
    for (int i = 0; i + 1 < N && rc == 0; ++i) {
        const int64_t nw2 = s->w[i] * d * d * d * d * s->w[i + 2];
        rc = alloc_d(&s->W2[i], nw2);
        if (rc == 0) rc = ct::heff_build_w2(h, CT_F64, s->W[i], s->W[i + 1], s->w[i], s->w[i + 1], s->w[i + 2], d, s->W2[i]);
        const int64_t th = s->Dcap[i] * d * d * s->Dcap[i + 2];
        if (th > theta_cap) theta_cap = th;
        const int64_t k = std::min(s->Dcap[i] * d, d * s->Dcap[i + 2]);
        if (k > s_cap) s_cap = k;
    }
    if (rc == 0) rc = alloc_d(&s->theta, theta_cap);
    if (rc == 0) rc = alloc_d(&s->S, s_cap);
    if (rc == 0) rc = ct::hip_rc(hipStreamSynchronize(h->stream));
    if (rc != 0) {
        delete s;
        return rc;
    }
    *out = s;
    return 0;
}

int ct_dmrg_destroy(ct_dmrg_t s) {
    delete s;
    return 0;
}

int ct_dmrg_run(ct_dmrg_t s, int num_sweeps, double* energy, double* sweep_energies, int* sweeps_run) {
    CT_ARG(s != nullptr, 1, "state is NULL");
    const int sweeps = num_sweeps > 0 ? num_sweeps : s->p.num_sweeps;
    const int N = s->N;
    const double t0 = now_s();
    if (sweeps_run) *sweeps_run = 0;
    CT_CHECK(build_envs(s));
    double E = 0.0, prev = 0.0;
    for (int sw = 0; sw < sweeps; ++sw) {
        for (int i = 0; i < N - 1; ++i) CT_CHECK(local_update(s, i, true, &E));
        for (int i = N - 2; i >= 0; --i) CT_CHECK(local_update(s, i, false, &E));
        if (sweep_energies) sweep_energies[sw] = E;
        s->sweep_E.push_back(E);
        s->st.sweeps += 1;
        if (sweeps_run) *sweeps_run = sw + 1;
        if (s->p.verbose) fprintf(stderr, "Sweep %d completed. Energy: %.15g\n", sw + 1, E);
        if (sw > 0 && std::fabs(E - prev) < s->p.energy_tol) break;
        prev = E;
    }
    CT_HIP(hipStreamSynchronize(s->h->stream));
    s->st.energy = E;
    s->st.t_total += now_s() - t0;
    if (energy) *energy = E;
    return 0;
}

int ct_dmrg_get_stats(ct_dmrg_t s, ct_dmrg_stats* stats) {
    CT_ARG(s != nullptr, 1, "state is NULL");
    CT_ARG(stats != nullptr, 2, "stats is NULL");
    *stats = s->st;
    return 0;
}

int ct_dmrg_get_dims(ct_dmrg_t s, int64_t* dims) {
    CT_ARG(s != nullptr, 1, "state is NULL");
    CT_ARG(dims != nullptr, 2, "dims is NULL");
    for (int l = 0; l <= s->N; ++l) dims[l] = s->D[l];
    return 0;
}

int ct_dmrg_get_site(ct_dmrg_t s, int i, int dtype, void* dst) {
    CT_ARG(s != nullptr, 1, "state is NULL");
    CT_ARG(i >= 0 && i < s->N, 2, "site index out of range");
    CT_ARG(dtype == CT_F64 || dtype == CT_F32, 3, "unknown dtype");
    const int64_t n = s->D[i] * s->d * s->D[i + 1];
    if (dtype == CT_F64)
        CT_HIP(hipMemcpyAsync(dst, s->A[i], (size_t)n * 8, hipMemcpyDeviceToDevice, s->h->stream));
    else
        CT_CHECK(ct::convert(s->h, CT_F64, s->A[i], CT_F32, dst, n));
    CT_HIP(hipStreamSynchronize(s->h->stream));
    return 0;
}

}  // extern "C"
