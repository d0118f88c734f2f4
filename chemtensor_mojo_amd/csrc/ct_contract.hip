// Contractions of the two-site DMRG hot path, composed from the MFMA GEMMs
// (ct_gemm*.hip) and the MPO mixing kernel (ct_ops.hip).
//
// apply_two_site_heff (reference src/algorithms/dmrg.mojo:254-374) keeps the
// reference contraction order L.theta -> W_i -> W_{i+1} -> R.  The two
// K = w*d MPO steps (:326-356) are fused into one pass with the pre-contracted
// coefficients W2 = W_i . W_{i+1} (exact reassociation), and the four
// physical permutes (:316, :331, :347, :363) disappear: the GEMM tile loaders
// read the operands in their stored layout and the mixing kernel writes its
// output directly in the layout the last GEMM consumes.
//
// HBM layout (row-major):
//   L [wL][Dl][Dlb], R [wR][Dr][Drb], theta/out [Dl][d][d][Dr]
//   T1 [wL][Dlb][d][d][Dr]    = L^T theta          (batched GEMM over wL)
//   U  [Dlb][d][d][wR][Dr]    = W2 mixing of T1    (per pixel (Dlb, Dr))
//   out[(Dlb d d)][Drb]       = U[(Dlb d d)][(wR Dr)] R[(wR Dr)][Drb]
#include "ct_internal.h"
#include "ct_device.h"

namespace ct {

namespace {

// coef for the fused two-site mixing; one thread per coefficient.
template <typename TW>
__global__ void build_w2_kernel(const TW* Wi, const TW* Wj, int wL, int wM, int wR, int d, double* out) {
    const int n = wL * d * d * d * d * wR;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    // t = ((wl*d + s1)*d + s2) * (wR*d*d) + (wr*d + s1o)*d + s2o
    const int NO = wR * d * d;
    int k = t / NO, c = t - k * NO;
    const int wl = k / (d * d), s1 = (k / d) % d, s2 = k % d;
    const int wr = c / (d * d), s1o = (c / d) % d, s2o = c % d;
    double acc = 0.0;
    for (int wm = 0; wm < wM; ++wm) {
        const double a = (double)Wi[((wl * d + s1) * d + s1o) * wM + wm];
        const double b = (double)Wj[((wm * d + s2) * d + s2o) * wR + wr];
        acc += a * b;
    }
    out[t] = acc;
}

// coef for the environment mixing.
//  left : coef[(wl*d + s)*(wR*d) + wr*d + so] = W[wl, s, so, wr]
//  right: coef[(wr*d + s)*(wL*d) + wl*d + so] = W[wl, s, so, wr]
template <typename TW>
__global__ void build_env_coef_kernel(const TW* W, int wL, int wR, int d, int right, double* out) {
    const int n = wL * d * d * wR;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    // t enumerates W[wl][s][so][wr]
    const int wr = t % wR, so = (t / wR) % d, s = (t / (wR * d)) % d, wl = t / (wR * d * d);
    const double v = (double)W[t];
    if (!right)
        out[(wl * d + s) * (wR * d) + wr * d + so] = v;
    else
        out[(wr * d + s) * (wL * d) + wl * d + so] = v;
}


}  // namespace

int heff_build_w2(Handle* h, int tW, const void* Wi, const void* Wj, int64_t wL, int64_t wM, int64_t wR, int64_t d,
                  double* d_out) {
    const int n = (int)(wL * d * d * d * d * wR);
    if (tW == CT_F64)
        hipLaunchKernelGGL(build_w2_kernel<double>, dim3((n + 255) / 256), dim3(256), 0, h->stream,
                           (const double*)Wi, (const double*)Wj, (int)wL, (int)wM, (int)wR, (int)d, d_out);
    else
        hipLaunchKernelGGL(build_w2_kernel<float>, dim3((n + 255) / 256), dim3(256), 0, h->stream, (const float*)Wi,
                           (const float*)Wj, (int)wL, (int)wM, (int)wR, (int)d, d_out);
    CT_HIP(hipGetLastError());
    return 0;
}

static int build_env_coef(Handle* h, int tW, const void* W, int64_t wL, int64_t wR, int64_t d, int right,
                          double* d_out) {
    const int n = (int)(wL * d * d * wR);
    if (tW == CT_F64)
        hipLaunchKernelGGL(build_env_coef_kernel<double>, dim3((n + 255) / 256), dim3(256), 0, h->stream,
                           (const double*)W, (int)wL, (int)wR, (int)d, right, d_out);
    else
        hipLaunchKernelGGL(build_env_coef_kernel<float>, dim3((n + 255) / 256), dim3(256), 0, h->stream,
                           (const float*)W, (int)wL, (int)wR, (int)d, right, d_out);
    CT_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------ H_eff
int heff_prepare(Handle* h, int prec, int tEnv, const void* L, const void* R, const double* d_W2,
                 const HeffDims& hd, HeffEnv* env, HeffStore* store) {
    env->prec = prec;
    env->W2 = d_W2;
    env->dims = hd;
    // workspace of the handle, or buffers of the prepared operator
    auto get = [&](int slot, Buffer* own, size_t bytes, void** out) {
        return store ? buf_get(h, *own, bytes, out) : ws_get(h, slot, bytes, out);
    };
    {   // padded mixing coefficients, shared by every matvec of this local problem
        const int tAcc = prec == CT_PREC_FP64 ? CT_F64 : CT_F32;
        const int K = (int)(hd.wL * hd.d * hd.d), NO = (int)(hd.d * hd.d * hd.wR);
        void* pad;
        CT_CHECK(get(tAcc == CT_F64 ? WS_W2PAD64 : WS_W2PAD32, store ? &store->pad : nullptr,
                     mix_pad_bytes(K, NO, tAcc), &pad));
        CT_CHECK(mix_pad_coef(h, d_W2, K, NO, tAcc, pad));
        env->W2pad = pad;
    }
    if (prec == CT_PREC_FP64) {
        if (tEnv != CT_F64) {
            set_last_error("FP64 contraction requires FP64 environments");
            return CT_STATUS_NOT_SUPPORTED;
        }
        env->tEnv = CT_F64;
        env->L = L;
        env->R = R;
        return 0;
    }
    if (prec == CT_PREC_FP32) {
        if (tEnv == CT_F32) {
            env->tEnv = CT_F32;
            env->L = L;
            env->R = R;
            return 0;
        }
        const int64_t nL = hd.wL * hd.Dl * hd.Dlb, nR = hd.wR * hd.Dr * hd.Drb;
        void *l32, *r32;
        CT_CHECK(get(WS_LSPLIT, store ? &store->l : nullptr, (size_t)nL * 4, &l32));
        CT_CHECK(get(WS_RSPLIT, store ? &store->r : nullptr, (size_t)nR * 4, &r32));
        CT_CHECK(convert(h, CT_F64, L, CT_F32, l32, nL));
        CT_CHECK(convert(h, CT_F64, R, CT_F32, r32, nR));
        env->tEnv = CT_F32;
        env->L = l32;
        env->R = r32;
        return 0;
    }
    if (prec == CT_PREC_BF16X3 || prec == CT_PREC_BF16X6) {
        const bool x6 = prec == CT_PREC_BF16X6;
        // Ls: split rows [wL][Dlb][k = Dl] from L[wL][Dl][Dlb]: row stride 1, k stride Dlb
        const int64_t KL32 = split_k32(hd.Dl), ldL = 2 * KL32, bsL = hd.Dlb * ldL, nL = hd.wL * bsL;
        // Rs: split rows [Drb][k = wR*Dr] from R[(wR Dr)][Drb]: row stride 1, k stride Drb
        const int64_t KR = hd.wR * hd.Dr, KR32 = split_k32(KR), ldR = 2 * KR32, nR = hd.Drb * ldR;
        void *ls, *rs, *ll = nullptr, *rl = nullptr;
        CT_CHECK(get(WS_LSPLIT, store ? &store->l : nullptr, (size_t)nL * 2, &ls));
        CT_CHECK(get(WS_RSPLIT, store ? &store->r : nullptr, (size_t)nR * 2, &rs));
        if (x6) {  // the l planes: half the split rows' size
            CT_CHECK(get(WS_LSPLIT_L, store ? &store->ll : nullptr, (size_t)nL, &ll));
            CT_CHECK(get(WS_RSPLIT_L, store ? &store->rl : nullptr, (size_t)nR, &rl));
        }
        CT_CHECK(split_bf16(h, tEnv, L, (int)hd.Dlb, (int)hd.Dl, 1, hd.Dlb, (int)hd.wL, hd.Dl * hd.Dlb,
                            (uint16_t*)ls, ldL, bsL, nullptr, (uint16_t*)ll, KL32, hd.Dlb * KL32));
        CT_CHECK(split_bf16(h, tEnv, R, (int)hd.Drb, (int)KR, 1, hd.Drb, 1, 0, (uint16_t*)rs, ldR, 0, nullptr,
                            (uint16_t*)rl, KR32, 0));
        env->tEnv = tEnv;
        env->L = L;
        env->R = R;
        env->Ls.p = (const uint16_t*)ls; env->Ls.ld = ldL; env->Ls.bs = bsL;
        env->Rs.p = (const uint16_t*)rs; env->Rs.ld = ldR; env->Rs.bs = 0;
        if (x6) {
            env->Ls.l = (const uint16_t*)ll; env->Ls.ldl = KL32; env->Ls.bsl = hd.Dlb * KL32;
            env->Rs.l = (const uint16_t*)rl; env->Rs.ldl = KR32; env->Rs.bsl = 0;
        }
        return 0;
    }
    set_last_error("unknown precision");
    return CT_STATUS_NOT_SUPPORTED;
}

int heff_apply(Handle* h, const HeffEnv& env, int tVec, const void* theta, void* out, const int* active) {
    const HeffDims& D = env.dims;
    const int64_t d2 = D.d * D.d;
    const int64_t nT1 = D.wL * D.Dlb * d2 * D.Dr;
    const int64_t nU = D.Dlb * d2 * D.wR * D.Dr;

    MixDesc m;
    m.CIN = (int)D.wL; m.EIN = (int)d2; m.COUT = (int)D.wR; m.EOUT = (int)d2;
    m.P = D.Dlb; m.Q = D.Dr;
    m.sXc = D.Dlb * d2 * D.Dr; m.sXp = d2 * D.Dr; m.sXe = D.Dr;
    m.sYc = D.Dr; m.sYp = d2 * D.wR * D.Dr; m.sYe = D.wR * D.Dr;
    m.active = active;

    if (env.prec == CT_PREC_BF16X3 || env.prec == CT_PREC_BF16X6) {
        const bool x6 = env.prec == CT_PREC_BF16X6;
        auto gemm_split = x6 ? gemm_bf16x6_nt : gemm_bf16x3_nt;
        // theta [Dl][(d d Dr)] -> split rows [n = d d Dr][k = Dl] (6xBF16: and its l plane)
        const int64_t Nn = d2 * D.Dr, KV32 = split_k32(D.Dl), ldV = 2 * KV32, nV = Nn * ldV;
        void *vs, *t1, *us, *vl = nullptr, *ul = nullptr;
        CT_CHECK(ws_get(h, WS_VSPLIT, (size_t)nV * 2, &vs));
        CT_CHECK(ws_get(h, WS_T1, (size_t)nT1 * 4, &t1));
        const int64_t KU = D.wR * D.Dr, KU32 = split_k32(KU), ldU = 2 * KU32, nUs = D.Dlb * d2 * ldU;
        CT_CHECK(ws_get(h, WS_USPLIT, (size_t)nUs * 2, &us));
        if (x6) {
            CT_CHECK(ws_get(h, WS_VSPLIT_L, (size_t)nV, &vl));
            CT_CHECK(ws_get(h, WS_USPLIT_L, (size_t)nUs, &ul));
        }
        CT_CHECK(split_bf16(h, tVec, theta, (int)Nn, (int)D.Dl, 1, Nn, 1, 0, (uint16_t*)vs, ldV, 0, active,
                            (uint16_t*)vl, KV32, 0));
        SplitOp sv;
        sv.p = (const uint16_t*)vs; sv.ld = ldV; sv.bs = 0;
        sv.l = (const uint16_t*)vl; sv.ldl = KV32; sv.bsl = 0;
        // T1[b][Dlb][(d d Dr)] = Ls[b] (Dlb x Dl) . Vs^T
        CT_CHECK(gemm_split(h, env.Ls, sv, CT_F32, t1, (int)D.Dlb, (int)Nn, (int)D.Dl, Nn, D.Dlb * Nn, (int)D.wL, 1.0,
                            0.0, active));
        // U written as split rows [(Dlb d d)][k = (wR Dr)] (6xBF16: and its l plane)
        uint16_t* uh = (uint16_t*)us;
        MixDesc ms = m;
        ms.sYc = D.Dr; ms.sYp = d2 * ldU; ms.sYe = ldU;
        CT_CHECK(split_zero_pad(h, uh, D.Dlb * d2, KU, ldU, (uint16_t*)ul, KU32));
        CT_CHECK(mix_split(h, (const float*)t1, uh, env.W2, ms, env.W2pad, (uint16_t*)ul));
        SplitOp su;
        su.p = uh; su.ld = ldU; su.bs = 0;
        su.l = (const uint16_t*)ul; su.ldl = KU32; su.bsl = 0;
        // out[(Dlb d d)][Drb] = Us . Rs^T
        return gemm_split(h, su, env.Rs, tVec, out, (int)(D.Dlb * d2), (int)D.Drb, (int)KU, D.Drb, 0, 1, 1.0, 0.0,
                          active);
    }

    const int tMid = env.prec == CT_PREC_FP64 ? CT_F64 : CT_F32;
    const size_t es = tMid == CT_F64 ? 8 : 4;
    void *t1, *u;
    CT_CHECK(ws_get(h, WS_T1, (size_t)nT1 * es, &t1));
    CT_CHECK(ws_get(h, WS_T2, (size_t)nU * es, &u));
    const void* vec = theta;
    int tV = tVec;
    if (env.prec == CT_PREC_FP32 && tVec == CT_F64) {
        // FP32 copy of the FP64 vector: step 1 reads theta once per MPO channel
        // (batch wL), and FP32 tiles load half the bytes with no conversions
        // (D = 1024, w = 8: step 1 ~600 -> ~540 us, for a ~10 us copy)
        void* v32;
        const int64_t nV = D.Dl * d2 * D.Dr;
        CT_CHECK(ws_get(h, WS_VSPLIT, (size_t)nV * 4, &v32));
        CT_CHECK(convert(h, CT_F64, theta, CT_F32, v32, nV, active));
        vec = v32;
        tV = CT_F32;
    }
    {   // step 1: T1[b][m=Dlb][n=(d d Dr)] = sum_k L[b][k][m] theta[k][n]
        GemmDesc g = heff_step1_desc(D);
        g.active = active;
        CT_CHECK(gemm(h, env.prec, env.tEnv, env.L, tV, vec, tMid, t1, g));
    }
    // steps 2+3: fused MPO mixing
    CT_CHECK(mix(h, tMid, t1, tMid, u, env.W2, m, env.W2pad));
    {   // step 4: out[(Dlb d d)][Drb] = U[(Dlb d d)][(wR Dr)] R[(wR Dr)][Drb]
        GemmDesc g = heff_step4_desc(D);
        g.active = active;
        CT_CHECK(gemm(h, env.prec, tMid, u, env.tEnv, env.R, tVec, out, g));
    }
    return 0;
}

GemmDesc heff_step1_desc(const HeffDims& D) {
    const int64_t d2 = D.d * D.d;
    GemmDesc g;
    g.M = (int)D.Dlb; g.N = (int)(d2 * D.Dr); g.K = (int)D.Dl; g.batch = (int)D.wL;
    g.a_kc = 0; g.sAm = 1; g.sAk = D.Dlb; g.bsA = D.Dl * D.Dlb;
    g.b_kc = 0; g.sBk = d2 * D.Dr; g.sBn = 1; g.bsB = 0;
    g.ldc = d2 * D.Dr; g.bsC = D.Dlb * d2 * D.Dr;
    return g;
}

GemmDesc heff_step4_desc(const HeffDims& D) {
    const int64_t d2 = D.d * D.d;
    GemmDesc g;
    g.M = (int)(D.Dlb * d2); g.N = (int)D.Drb; g.K = (int)(D.wR * D.Dr); g.batch = 1;
    g.a_kc = 1; g.sAm = D.wR * D.Dr; g.sAk = 1;
    g.b_kc = 0; g.sBk = D.Drb; g.sBn = 1;
    g.ldc = D.Drb;
    return g;
}

// ------------------------------------------------------------ environments
// update_left_environment (src/state/environments.mojo:99-201), same three
// contractions; the four permutes are folded into the operand layouts.
//   T1[wL][Dl'][s][Dr] = sum_Dl L[wL][Dl][Dl'] A[Dl][(s Dr)]
//   V [wR][Dl'][s'][Dr] = sum_{wL,s} W[wL,s,s',wR] T1[wL][Dl'][s][Dr]
//   out[wR][Dr][Dr']    = sum_{(Dl' s')} V[wR][(Dl' s')][Dr] A[(Dl' s')][Dr']
int env_left(Handle* h, int prec, int dtype, const void* L, const void* A, const void* W, void* out, int64_t Dl,
             int64_t d, int64_t Dr, int64_t wL, int64_t wR) {
    const int tMid = prec == CT_PREC_FP64 ? CT_F64 : CT_F32;
    const size_t es = tMid == CT_F64 ? 8 : 4;
    const int64_t n1 = wL * Dl * d * Dr, n2 = wR * Dl * d * Dr;
    void *t1, *v, *coef;
    CT_CHECK(ws_get(h, WS_ENV_T1, (size_t)n1 * es, &t1));
    CT_CHECK(ws_get(h, WS_ENV_T2, (size_t)n2 * es, &v));
    CT_CHECK(ws_get(h, WS_ENV_CONV, (size_t)(wL * d * d * wR) * 8, &coef));
    CT_CHECK(build_env_coef(h, dtype, W, wL, wR, d, 0, (double*)coef));
    {
        GemmDesc g;
        g.M = (int)Dl; g.N = (int)(d * Dr); g.K = (int)Dl; g.batch = (int)wL;
        g.a_kc = 0; g.sAm = 1; g.sAk = Dl; g.bsA = Dl * Dl;
        g.b_kc = 0; g.sBk = d * Dr; g.sBn = 1; g.bsB = 0;
        g.ldc = d * Dr; g.bsC = Dl * d * Dr;
        CT_CHECK(gemm(h, prec, dtype, L, dtype, A, tMid, t1, g));
    }
    {
        MixDesc m;
        m.CIN = (int)wL; m.EIN = (int)d; m.COUT = (int)wR; m.EOUT = (int)d;
        m.P = Dl; m.Q = Dr;
        m.sXc = Dl * d * Dr; m.sXp = d * Dr; m.sXe = Dr;
        m.sYc = Dl * d * Dr; m.sYp = d * Dr; m.sYe = Dr;
        CT_CHECK(mix(h, tMid, t1, tMid, v, (const double*)coef, m));
    }
    {
        GemmDesc g;
        g.M = (int)Dr; g.N = (int)Dr; g.K = (int)(Dl * d); g.batch = (int)wR;
        g.a_kc = 0; g.sAm = 1; g.sAk = Dr; g.bsA = Dl * d * Dr;
        g.b_kc = 0; g.sBk = Dr; g.sBn = 1; g.bsB = 0;
        g.ldc = Dr; g.bsC = Dr * Dr;
        CT_CHECK(gemm(h, prec, tMid, v, dtype, A, dtype, out, g));
    }
    return 0;
}

// update_right_environment (src/state/environments.mojo:309-417):
//   T1[wR][(Dl s)][Dr'] = sum_Dr A[(Dl s)][Dr] R[wR][Dr][Dr']
//   V [wL][Dl][s'][Dr'] = sum_{wR,s} W[wL,s,s',wR] T1[wR][Dl][s][Dr']
//   out[wL][Dl][Dl']    = sum_{(s' Dr')} V[wL][Dl][(s' Dr')] A[Dl'][(s' Dr')]
int env_right(Handle* h, int prec, int dtype, const void* R, const void* A, const void* W, void* out, int64_t Dl,
              int64_t d, int64_t Dr, int64_t wL, int64_t wR) {
    const int tMid = prec == CT_PREC_FP64 ? CT_F64 : CT_F32;
    const size_t es = tMid == CT_F64 ? 8 : 4;
    const int64_t n1 = wR * Dl * d * Dr, n2 = wL * Dl * d * Dr;
    void *t1, *v, *coef;
    CT_CHECK(ws_get(h, WS_ENV_T1, (size_t)n1 * es, &t1));
    CT_CHECK(ws_get(h, WS_ENV_T2, (size_t)n2 * es, &v));
    CT_CHECK(ws_get(h, WS_ENV_CONV, (size_t)(wL * d * d * wR) * 8, &coef));
    CT_CHECK(build_env_coef(h, dtype, W, wL, wR, d, 1, (double*)coef));
    {
        GemmDesc g;
        g.M = (int)(Dl * d); g.N = (int)Dr; g.K = (int)Dr; g.batch = (int)wR;
        g.a_kc = 1; g.sAm = Dr; g.sAk = 1; g.bsA = 0;
        g.b_kc = 0; g.sBk = Dr; g.sBn = 1; g.bsB = Dr * Dr;
        g.ldc = Dr; g.bsC = Dl * d * Dr;
        CT_CHECK(gemm(h, prec, dtype, A, dtype, R, tMid, t1, g));
    }
    {
        MixDesc m;
        m.CIN = (int)wR; m.EIN = (int)d; m.COUT = (int)wL; m.EOUT = (int)d;
        m.P = Dl; m.Q = Dr;
        m.sXc = Dl * d * Dr; m.sXp = d * Dr; m.sXe = Dr;
        m.sYc = Dl * d * Dr; m.sYp = d * Dr; m.sYe = Dr;
        CT_CHECK(mix(h, tMid, t1, tMid, v, (const double*)coef, m));
    }
    {
        GemmDesc g;
        g.M = (int)Dl; g.N = (int)Dl; g.K = (int)(d * Dr); g.batch = (int)wL;
        g.a_kc = 1; g.sAm = d * Dr; g.sAk = 1; g.bsA = Dl * d * Dr;
        g.b_kc = 1; g.sBn = d * Dr; g.sBk = 1; g.bsB = 0;
        g.ldc = Dl; g.bsC = Dl * Dl;
        CT_CHECK(gemm(h, prec, tMid, v, dtype, A, dtype, out, g));
    }
    return 0;
}

// --------------------------------------------------------------- tensordot
// dense_tensor_dot (src/m_tensor/dense_tensor.mojo:963-1331).  Axis
// inference (:1141-1159), shape checks (:1161-1167, :1283-1299), then one
// GEMM whose loaders read A as [m][k] or [k][m] and B as [k][n] or [n][k]
// directly (the reference's physical transposes :1249-1257 are not needed).
int tensordot(Handle* h, int prec, int dtype, const void* A, int rA, const int64_t* sA, const void* B, int rB,
              const int64_t* sB, int nm, int axA, int axB, void* C, int rC, const int64_t* sC) {
    CT_ARG(rA >= 1 && rA <= 16, 5, "rankA out of range");
    CT_ARG(rB >= 1 && rB <= 16, 8, "rankB out of range");
    if (nm < 1 || rA < nm || rB < nm) {
        set_last_error("Invalid ndim_mult");
        return -10;
    }
    int effB = axB;
    if (!axA && !axB) {
        bool lead = true, trail = true;
        for (int i = 0; i < nm; ++i) {
            const int64_t a = sA[rA - nm + i];
            if (a != sB[i]) lead = false;
            if (a != sB[rB - nm + i]) trail = false;
        }
        effB = lead ? 1 : (trail ? 0 : 1);
    }
    for (int i = 0; i < nm; ++i) {
        if (sA[(axA ? 0 : rA - nm) + i] != sB[(effB ? 0 : rB - nm) + i]) {
            set_last_error("Dim mismatch");
            return CT_STATUS_SHAPE;
        }
    }
    std::vector<int64_t> expect;
    for (int i = axA ? nm : 0; i < (axA ? rA : rA - nm); ++i) expect.push_back(sA[i]);
    for (int i = effB ? nm : 0; i < (effB ? rB : rB - nm); ++i) expect.push_back(sB[i]);
    if (expect.empty()) {
        if (!(rC == 0 || (rC == 1 && sC[0] == 1))) {
            set_last_error("Output tensor C for scalar result must be [] or [1]");
            return CT_STATUS_SHAPE;
        }
    } else {
        if ((size_t)rC != expect.size()) {
            set_last_error("Output tensor C has wrong rank");
            return CT_STATUS_SHAPE;
        }
        for (int i = 0; i < rC; ++i)
            if (sC[i] != expect[i]) {
                set_last_error("Output tensor C has wrong shape");
                return CT_STATUS_SHAPE;
            }
    }
    int64_t k = 1, m = 1, n = 1;
    for (int i = 0; i < nm; ++i) k *= sA[(axA ? 0 : rA - nm) + i];
    for (int i = axA ? nm : 0; i < (axA ? rA : rA - nm); ++i) m *= sA[i];
    for (int i = effB ? nm : 0; i < (effB ? rB : rB - nm); ++i) n *= sB[i];
    GemmDesc g;
    g.M = (int)m; g.N = (int)n; g.K = (int)k; g.batch = 1;
    if (axA) { g.a_kc = 0; g.sAm = 1; g.sAk = m; }
    else { g.a_kc = 1; g.sAm = k; g.sAk = 1; }
    if (effB) { g.b_kc = 0; g.sBk = n; g.sBn = 1; }
    else { g.b_kc = 1; g.sBn = k; g.sBk = 1; }
    g.ldc = n;
    return gemm(h, prec, dtype, A, dtype, B, dtype, C, g);
}

// ----------------------------------------------------------------- adjoint
// gWi[wl,s1,s1',wm] = sum L[wl,a,a'] th[a,s1,s2,b] Wj[wm,s2,s2',wr] R[wr,b,b'] th[a',s1',s2',b']
// gWj[wm,s2,s2',wr] = sum L th Wi R th
// Computed as (e = (s1 s2), eo = (s1' s2'), E = d*d)
//   T1[wl][e][a'][b]  = sum_a L[wl][a][a'] th[a][e][b]             (E GEMMs, batched over wl)
//   X [wr][eo][a'][b] = sum_b' th[a'][eo][b'] R[wr][b][b']          (E GEMMs, batched over wr)
//   G [wl,e ; wr,eo]  = sum_{(a',b)} T1[(wl,e)][(a',b)] X[(wr,eo)][(a',b)]
// The (e, a', b) order makes (a', b) one contiguous index of length Dlb*Dr, so
// G is a single k-contiguous GEMM of M = wL*E, N = wR*E, K = Dlb*Dr, split over
// K across the CUs: T1 and X are each read once.  Then the tiny contractions of
// G with Wj (for gWi) or Wi (for gWj).
namespace {
__global__ void adj_contract_kernel(const double* __restrict__ G, const double* __restrict__ Wo, int wL, int wM,
                                    int wR, int d, int which, double* __restrict__ out) {
    // G[((wl*d+s1)*d+s2) * (wR*d*d) + (wr*d+s1o)*d+s2o]
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int NO = wR * d * d;
    if (which == 0) {
        // out = gWi[wl][s1][s1o][wm] = sum_{s2,s2o,wr} G[wl,s1,s2;wr,s1o,s2o] Wj[wm,s2,s2o,wr]
        const int n = wL * d * d * wM;
        if (t >= n) return;
        const int wm = t % wM, s1o = (t / wM) % d, s1 = (t / (wM * d)) % d, wl = t / (wM * d * d);
        double acc = 0.0;
        for (int s2 = 0; s2 < d; ++s2)
            for (int s2o = 0; s2o < d; ++s2o)
                for (int wr = 0; wr < wR; ++wr)
                    acc += G[((wl * d + s1) * d + s2) * NO + (wr * d + s1o) * d + s2o] *
                           Wo[((wm * d + s2) * d + s2o) * wR + wr];
        out[t] = acc;
    } else {
        // out = gWj[wm][s2][s2o][wr] = sum_{wl,s1,s1o} G[wl,s1,s2;wr,s1o,s2o] Wi[wl,s1,s1o,wm]
        const int n = wM * d * d * wR;
        if (t >= n) return;
        const int wr = t % wR, s2o = (t / wR) % d, s2 = (t / (wR * d)) % d, wm = t / (wR * d * d);
        double acc = 0.0;
        for (int wl = 0; wl < wL; ++wl)
            for (int s1 = 0; s1 < d; ++s1)
                for (int s1o = 0; s1o < d; ++s1o)
                    acc += G[((wl * d + s1) * d + s2) * NO + (wr * d + s1o) * d + s2o] *
                           Wo[((wl * d + s1) * d + s1o) * wM + wm];
        out[t] = acc;
    }
}

// G[(wl, e) ; (wr, eo)] = sum_{(a, b)} T1[wl][e][(a, b)] * X[wr][eo][(a, b)]:
// one k-contiguous GEMM, C = A B^T, with K = P*Q (both operands stored [w][e][P*Q]).
int adj_gram(Handle* h, const double* T1, const double* X, int64_t P, int64_t Q, int64_t E, int64_t wL, int64_t wR,
             double* G) {
    if (P * Q > INT32_MAX || wL * E > INT32_MAX || wR * E > INT32_MAX) {
        set_last_error("adjoint: Dl*Dr or w*E exceeds the 32-bit GEMM extents");
        return CT_STATUS_NOT_SUPPORTED;
    }
    GemmDesc q;
    q.M = (int)(wL * E); q.N = (int)(wR * E); q.K = (int)(P * Q); q.batch = 1;
    q.a_kc = 1; q.sAm = P * Q; q.sAk = 1;
    q.b_kc = 1; q.sBn = P * Q; q.sBk = 1;
    q.ldc = wR * E;
    return gemm(h, CT_PREC_FP64, CT_F64, T1, CT_F64, X, CT_F64, G, q);
}
}  // namespace

int heff_adjoint(Handle* h, const double* L, const double* Wi, const double* Wj, const double* R,
                 const double* theta, double* gWi, double* gWj, const HeffDims& D) {
    const int64_t d2 = D.d * D.d;
    const int64_t nT1 = D.wL * D.Dlb * d2 * D.Dr, nX = D.wR * D.Dr * D.Dlb * d2;
    const int64_t nG = D.wL * d2 * D.wR * d2;
    void *t1, *x, *g;
    CT_CHECK(ws_get(h, WS_TMP, (size_t)nT1 * 8, &t1));
    CT_CHECK(ws_get(h, WS_TMP2, (size_t)nX * 8, &x));
    CT_CHECK(ws_get(h, WS_TMP3, (size_t)nG * 8, &g));
    for (int64_t e = 0; e < d2; ++e) {  // T1[wl][e][a'][b] = sum_a L[wl][a][a'] th[a][e][b]
        GemmDesc q;
        q.M = (int)D.Dlb; q.N = (int)D.Dr; q.K = (int)D.Dl; q.batch = (int)D.wL;
        q.a_kc = 0; q.sAm = 1; q.sAk = D.Dlb; q.bsA = D.Dl * D.Dlb;
        q.b_kc = 0; q.sBk = d2 * D.Dr; q.sBn = 1; q.bsB = 0;
        q.ldc = D.Dr; q.bsC = d2 * D.Dlb * D.Dr;
        CT_CHECK(gemm(h, CT_PREC_FP64, CT_F64, L, CT_F64, theta + e * D.Dr, CT_F64, (double*)t1 + e * D.Dlb * D.Dr, q));
    }
    for (int64_t e = 0; e < d2; ++e) {  // X[wr][eo][a'][b] = sum_{b'} th[a'][eo][b'] R[wr][b][b']
        GemmDesc q;
        q.M = (int)D.Dlb; q.N = (int)D.Dr; q.K = (int)D.Drb; q.batch = (int)D.wR;
        q.a_kc = 1; q.sAm = d2 * D.Drb; q.sAk = 1; q.bsA = 0;
        q.b_kc = 1; q.sBn = D.Drb; q.sBk = 1; q.bsB = D.Dr * D.Drb;
        q.ldc = D.Dr; q.bsC = d2 * D.Dlb * D.Dr;
        CT_CHECK(gemm(h, CT_PREC_FP64, CT_F64, theta + e * D.Drb, CT_F64, R, CT_F64, (double*)x + e * D.Dlb * D.Dr, q));
    }
    CT_CHECK(adj_gram(h, (const double*)t1, (const double*)x, D.Dlb, D.Dr, d2, D.wL, D.wR, (double*)g));
    if (gWi) {
        const int n = (int)(D.wL * d2 * D.wM);
        hipLaunchKernelGGL(adj_contract_kernel, dim3((n + 255) / 256), dim3(256), 0, h->stream, (const double*)g, Wj,
                           (int)D.wL, (int)D.wM, (int)D.wR, (int)D.d, 0, gWi);
        CT_HIP(hipGetLastError());
    }
    if (gWj) {
        const int n = (int)(D.wM * d2 * D.wR);
        hipLaunchKernelGGL(adj_contract_kernel, dim3((n + 255) / 256), dim3(256), 0, h->stream, (const double*)g, Wi,
                           (int)D.wL, (int)D.wM, (int)D.wR, (int)D.d, 1, gWj);
        CT_HIP(hipGetLastError());
    }
    return 0;
}

// Single-site environment adjoint:
//   gW[wl,s,s',wr] = sum L[wl,a,a'] A[a,s,b] A[a',s',b'] R[wr,b,b']
// the derivative of <psi|H|psi> = L_i . A W A . R_{i+1} with respect to W.
//   T1[wl][s][a'][b]  = sum_a L[wl][a][a'] A[a][s][b]       (d GEMMs batched over wl)
//   Y [wr][s'][a'][b] = sum_b' A[a'][s'][b'] R[wr][b][b']   (d GEMMs batched over wr)
//   G [wl,s ; wr,s']  = sum_{(a',b)} T1 Y                   (adj_gram, E = d)
// and G[wl][s][wr][s'] permuted to gW[wl][s][s'][wr].
int env_adjoint(Handle* h, const double* L, const double* A, const double* R, double* gW, int64_t Dl, int64_t d,
                int64_t Dr, int64_t wL, int64_t wR) {
    const int64_t nT1 = wL * Dl * d * Dr, nY = wR * Dl * d * Dr, nG = wL * d * d * wR;
    void *t1, *y;
    CT_CHECK(ws_get(h, WS_TMP, (size_t)nT1 * 8, &t1));
    CT_CHECK(ws_get(h, WS_TMP2, (size_t)nY * 8, &y));
    for (int64_t s = 0; s < d; ++s) {
        GemmDesc q;
        q.M = (int)Dl; q.N = (int)Dr; q.K = (int)Dl; q.batch = (int)wL;
        q.a_kc = 0; q.sAm = 1; q.sAk = Dl; q.bsA = Dl * Dl;
        q.b_kc = 0; q.sBk = d * Dr; q.sBn = 1; q.bsB = 0;
        q.ldc = Dr; q.bsC = d * Dl * Dr;
        CT_CHECK(gemm(h, CT_PREC_FP64, CT_F64, L, CT_F64, A + s * Dr, CT_F64, (double*)t1 + s * Dl * Dr, q));
    }
    for (int64_t s = 0; s < d; ++s) {
        GemmDesc q;
        q.M = (int)Dl; q.N = (int)Dr; q.K = (int)Dr; q.batch = (int)wR;
        q.a_kc = 1; q.sAm = d * Dr; q.sAk = 1; q.bsA = 0;
        q.b_kc = 1; q.sBn = Dr; q.sBk = 1; q.bsB = Dr * Dr;
        q.ldc = Dr; q.bsC = d * Dl * Dr;
        CT_CHECK(gemm(h, CT_PREC_FP64, CT_F64, A + s * Dr, CT_F64, R, CT_F64, (double*)y + s * Dl * Dr, q));
    }
    void* g;
    CT_CHECK(ws_get(h, WS_TMP3, (size_t)nG * 8, &g));
    CT_CHECK(adj_gram(h, (const double*)t1, (const double*)y, Dl, Dr, d, wL, wR, (double*)g));
    // G[wl][s][wr][s'] -> gW[wl][s][s'][wr]
    const int64_t shape[4] = {wL, d, wR, d};
    const int perm[4] = {0, 1, 3, 2};
    return permute(h, CT_F64, g, 4, shape, perm, gW);
}

}  // namespace ct
