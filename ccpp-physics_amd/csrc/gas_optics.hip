// k-distribution gas optics on gfx950.
//
// Restates, in a layout of its own, what ty_gas_optics_rrtmgp%gas_optics does
// inside rte-rrtmgp (called at physics/Radiation/RRTMGP/rrtmgp_lw_main.F90:337-345
// and rrtmgp_sw_main.F90:346-353): get_col_dry, interpolation,
// gas_optical_depths_major (interpolate3D), gas_optical_depths_minor
// (interpolate2D), compute_Planck_source's Planck-fraction interpolation,
// and for SW compute_tau_rayleigh.
//
// One kernel, gas_optics_kernel: block = (16-layer chunk, column), lane =
// g-point.
//   1. Interpolation (FP64), one thread per layer of the chunk: col_dry and
//      the gas columns, jtemp/ftemp, jpress/fpress, the tropopause flag, and
//      per (flavor, T-corner) col_mix, jeta and the fminor weights, per
//      minor interval the absorber scaling.  The discrete decisions
//      (indices, tropopause) are taken in FP64 as in the FP64 reference.
//      The result is one compact record per layer (layout below), built in
//      LDS; it never goes to HBM.
//   2. Per (layer, item) table offsets and weights, built once per block.
//   3. Gathers (FP32): 8 major-table corners (+8 Planck-fraction corners LW,
//      +4 Rayleigh corners SW) and the minor intervals covering the g-point,
//      two layers' gathers in flight.
// Blocks are numbered so that each of the 8 XCDs (which take workgroups in
// turn) walks one contiguous range of (chunk, column): its L2 holds the
// k-table rows of a few chunks only, and the CCPP input lines of a run of
// neighbouring columns.
// Output tau / pfrac / tau_ray : [col][lay][gpt] FP32.
//
// Per-layer record in LDS (32-bit words), NMN = the larger regime's minor
// count:
//   [0]       jtemp | jpress << 8 | itropo << 16
//   [1]       fpress  [2] cray (column of dry air + water vapour, Rayleigh)
//   [3 + m]   scaling of minor interval m of the layer's regime;
//             [3 + NMN] = 0, the scale of unused item slots
// and per (layer, item), item = 2 flavor + T-corner, two quads of gather
// constants (interp_item).
#include "kernels.h"

namespace ccpprad {

constexpr int LCHUNK = 16;

// record geometry
struct RecGeom {
    int nf2, ms, nmn, words;
};
__host__ __device__ inline RecGeom rec_geom(int nflav, int nminor0, int nminor1)
{
    RecGeom r;
    r.nf2 = 2 * nflav;
    r.ms = 3;
    r.nmn = nminor0 > nminor1 ? nminor0 : nminor1;
    r.words = 4 * ((r.ms + r.nmn + 1 + 3) / 4);
    return r;
}

// ---------------------------------------------------------------------------
// interpolation, FP64
// ---------------------------------------------------------------------------
// v: the lane's gas columns in LDS (index 0 = dry air; a register array
// indexed at run time would go to scratch); i outside 1..n-1 reads the
// dry-air column
__device__ __forceinline__ double pick(const double *v, int n, int i)
{
    return v[(i >= 1 && i < n) ? i : 0];
}


// Per-layer FP64 scratch of the interpolation in LDS: the gas columns
// (index 0 = dry air), then ftemp, the density factor 0.01 p / T of the
// minor scalings, 1 / col_dry and 1 / (1 + col_h2o / col_dry).
constexpr int ROW_DBL = MAX_GAS + 1 + 4;

// Interpolation, part 1 (one thread per layer): get_col_dry, the
// temperature and pressure indices and weights, the tropopause flag and the
// Rayleigh column; record words 0-2 and the layer's scratch.
__device__ __forceinline__ void interp_layer(const TablesView &T, int l, int d, int ncol_ld, const int *col_idx,
                                             int col0, const double *p_lay, const double *p_lev, const double *t_lay,
                                             const VmrPtrs &vmr, unsigned *r, double *rs, int *err)
{
    const int gc = col_idx ? col_idx[col0 + d] - 1 : col0 + d;
    const size_t q = (size_t)gc + (size_t)ncol_ld * l;
    const double tmin = T.temp_ref_min, dT = T.temp_ref_delta;
    const double pl = p_lay[q], tl = t_lay[q];
    if (!(pl >= T.press_min && pl <= T.press_max && tl >= T.temp_min && tl <= T.temp_max)) atomicOr(err, 1);
    const double dp = fabs(p_lev[q] - p_lev[q + ncol_ld]);
    const double vh = vmr.v[0][q];  // gas 1 is h2o
    const double fact = 1.0 / (1.0 + vh);
    const double m_air = (M_DRY + M_H2O * vh) * fact;
    const double col_dry = 10.0 * dp * AVOGAD / (1000.0 * m_air * 100.0 * GRAV) * fact;
    rs[0] = col_dry;
#pragma unroll
    for (int gs = 1; gs <= MAX_GAS; ++gs) rs[gs] = gs <= T.ngas ? vmr.v[gs - 1][q] * col_dry : 0.0;
    int jt = (int)((tl - (tmin - dT)) / dT);
    jt = min(T.ntemp - 1, max(1, jt));
    const double lpl = log(pl);
    const double locp = 1.0 + (T.press_ref_log1 - lpl) / T.press_ref_log_delta;
    int jp = (int)locp;
    jp = min(T.npres - 1, max(1, jp));
    const int itr = (lpl > T.press_ref_trop_log) ? 1 : 2;
    rs[MAX_GAS + 1] = (tl - T.temp_ref[jt - 1]) / dT;     // ftemp
    rs[MAX_GAS + 2] = 0.01 * pl / tl;                      // density factor
    rs[MAX_GAS + 3] = 1.0 / col_dry;
    rs[MAX_GAS + 4] = 1.0 / (1.0 + rs[1] * rs[MAX_GAS + 3]);
    r[0] = (unsigned)jt | ((unsigned)jp << 8) | ((unsigned)itr << 16);
    r[1] = __float_as_uint((float)(locp - (double)jp));
    r[2] = __float_as_uint((float)(rs[1] + col_dry));
}

// Interpolation, part 2 (one thread per (layer, item), item = 2 flavor +
// T-corner): col_mix, jeta and the eta / temperature weights, turned
// straight into the gather constants of the item: quad 0 the major-table
// element offset of the (jtemp + t, jpress, jeta) corner without the
// g-point, fe0, fe1 (eta weights times the temperature weight), col_mix;
// quad 1 (jtemp + t) * neta + jeta - 1 (row of the minor and Rayleigh
// tables), fe0, fe1, 0.
__device__ __forceinline__ void interp_item(const TablesView &T, const unsigned *r, const double *rs, int it,
                                            uint4 *item)
{
    const int ng1 = T.ngas + 1, f = it >> 1, t = it & 1;
    const unsigned w0 = r[0];
    const int jt = (int)(w0 & 0xFFu), jp = (int)((w0 >> 8) & 0xFFu), itr = (int)(w0 >> 16);
    const double ftemp = rs[MAX_GAS + 1];
    const int g1 = T.flavor[2 * f], g2 = T.flavor[2 * f + 1];
    const double c1 = pick(rs, ng1, g1), c2 = pick(rs, ng1, g2);
    const double ratio = T.flav_ratio[((size_t)(jt - 1 + t) * T.nflav + f) * 2 + (itr - 1)];
    const double cm = c1 + ratio * c2;
    const double eta = (cm > 2.0 * 2.2250738585072014e-308) ? c1 / cm : 0.5;
    const double loceta = eta * (double)(T.neta - 1);
    const int jeta = min((int)loceta + 1, T.neta - 1);
    const double feta = loceta - floor(loceta);   // fmod(loceta, 1) for loceta >= 0
    const double ftt = t == 0 ? (1.0 - ftemp) : ftemp;
    const unsigned fe0 = __float_as_uint((float)((1.0 - feta) * ftt));
    const unsigned fe1 = __float_as_uint((float)(feta * ftt));
    const unsigned jt0 = (unsigned)(jt - 1 + t), je0 = (unsigned)(jeta - 1);
    const unsigned jp0 = (unsigned)(jp + itr - 2);   // 0-based lower pressure row
    const unsigned neta = (unsigned)T.neta, np1 = (unsigned)(T.npres + 1), ngpt = (unsigned)T.ngpt;
    const unsigned ko = __umul24(__umul24(__umul24(jt0, np1) + jp0, neta) + je0, ngpt);
    item[0] = make_uint4(ko, fe0, fe1, __float_as_uint((float)cm));
    item[1] = make_uint4(__umul24(jt0, neta) + je0, fe0, fe1, 0u);
}

// Interpolation, part 3 (one thread per (layer, minor interval)): the
// absorber scaling of minor interval m of the layer's regime (0 for m past
// the regime's count), record word MS + m.
__device__ __forceinline__ void interp_minor(const TablesView &T, const RecGeom &G, unsigned *r, const double *rs,
                                             int m)
{
    const int reg = (int)(r[0] >> 16) - 1, ng1 = T.ngas + 1;
    double sc = 0.0;
    if (m < T.nminor[reg]) {
        const int *par = T.minor_par[reg] + 5 * m;
        sc = pick(rs, ng1, par[0]);
        if (par[1]) {
            sc *= rs[MAX_GAS + 2];
            if (par[2] > 0) {
                const double x = pick(rs, ng1, par[2]) * rs[MAX_GAS + 3] * rs[MAX_GAS + 4];
                sc *= par[3] ? (1.0 - x) : x;
            }
        }
    }
    r[G.ms + m] = __float_as_uint((float)sc);
}

// ---------------------------------------------------------------------------
// g-point gathers: block = (chunk, column), lane = g-point, FP32
// ---------------------------------------------------------------------------

// A minor absorber item of one g-point: contributor column (16 bits),
// interval (8 bits), flavor (8 bits).
__device__ __forceinline__ int mi_ci(unsigned v) { return (int)(v & 0xFFFFu); }
__device__ __forceinline__ int mi_iv(unsigned v) { return (int)((v >> 16) & 0xFFu); }
__device__ __forceinline__ int mi_fv(unsigned v) { return (int)(v >> 24); }

// Gathers through a uniform table base and a 32-bit element offset, as global
// loads: a generic (flat) load would also count against the LDS counter, so
// every LDS read of the record would wait for the table loads in flight.
typedef const __attribute__((address_space(1))) char gchar;
__device__ __forceinline__ float ld1(const float *base, unsigned idx)
{
    return *(const __attribute__((address_space(1))) float *)((gchar *)base + (idx << 2));
}
typedef float v2f __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float2 ld2(const float2 *base, unsigned idx)
{
    const v2f v = *(const __attribute__((address_space(1))) v2f *)((gchar *)base + (idx << 3));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 ld4(const float4 *base, unsigned idx)
{
    const v4f v = *(const __attribute__((address_space(1))) v4f *)((gchar *)base + (idx << 4));
    return make_float4(v.x, v.y, v.z, v.w);
}

// Per (layer, item) constants of a chunk, built once per block so that the
// lanes of a flavor do not each redo the index arithmetic:
//   quad 0: major-table element offset of the (jtemp + t, jpress, jeta)
//           corner without the g-point, fe0, fe1 (eta weights times the
//           temperature weight), col_mix
//   quad 1: (jtemp + t) * neta + jeta - 1 (row of the minor and Rayleigh
//           tables), fe0, fe1, 0
__host__ __device__ inline size_t gather_smem(const RecGeom &G)
{
    return (size_t)LCHUNK * (sizeof(unsigned) * G.words + 2 * sizeof(uint4) * G.nf2 + sizeof(double) * ROW_DBL);
}

size_t gas_gather_smem(int is_lw, int nflav, int nminor0, int nminor1, int ngpt, int neta)
{
    (void)is_lw; (void)ngpt; (void)neta;
    return gather_smem(rec_geom(nflav, nminor0, nminor1));
}

// layers whose gathers are in flight together (3 or 4 measured slower for
// both LW and SW: fewer wavefronts fit)
template <bool LW> constexpr int gather_layers() { return 2; }

// The block's (chunk, column): workgroup b runs on XCD b mod 8, which takes
// the b / 8-th item of its own contiguous eighth of the nchunk * ncols
// (chunk, column) pairs, column fastest.  Returns false for the few padding
// blocks past the end.
__device__ __forceinline__ bool block_item(int ncols, int nchunk, int &chunk, int &d)
{
    const unsigned nb = (unsigned)ncols * (unsigned)nchunk;
    const unsigned per = (nb + 7u) >> 3;
    const unsigned b = blockIdx.x;
    const unsigned j = (b & 7u) * per + (b >> 3);
    if (j >= nb) return false;
    chunk = (int)(j / (unsigned)ncols);
    d = (int)(j - (unsigned)chunk * (unsigned)ncols);
    return true;
}

// FB: the emulator's fallback launch (only rows with row_mask set, of
// columns with col_mask set), a separate instantiation so that profiles
// show it apart from the table path.
// At most 72 VGPRs, 7 wavefronts per SIMD, as many as the block's LDS allows
// with 9 flavors: without the bound the LW kernel used 73 (allocated in
// steps of 8: 80), which left 6 per SIMD (LW gas optics 27.3 -> 26.2 ms per
// 262,144-column step).
template <bool LW, int NM, bool FB>
__global__ __launch_bounds__(128, 7) void gas_optics_kernel(
    TablesView T, int nlay, int ncol_ld, const int *__restrict__ col_idx, int col0, int ncols,
    const double *__restrict__ p_lay, const double *__restrict__ p_lev, const double *__restrict__ t_lay,
    VmrPtrs vmr, int *__restrict__ err, float *__restrict__ tau,
    float *__restrict__ aux /* LW pfrac, SW tau_ray */, const unsigned char *__restrict__ row_mask,
    const unsigned char *__restrict__ col_mask)
{
    extern __shared__ uint4 smem_q[];
    const int nchunk = (nlay + LCHUNK - 1) / LCHUNK;
    int chunk, d;   // chunk of layers, column within this launch
    if (!block_item(ncols, nchunk, chunk, d)) return;
    const int tid = threadIdx.x;
    if (FB && !col_mask[d]) return;            // uniform per block
    const int l0 = chunk * LCHUNK;
    const int nl = min(LCHUNK, nlay - l0);
    if (FB) {
        // only chunks holding a flagged layer do any work
        bool any = false;
        for (int l = l0; l < l0 + nl; ++l) any |= row_mask[(size_t)d * nlay + l] != 0;
        if (!any) return;                      // uniform per block
    }
    const RecGeom G = rec_geom(T.nflav, T.nminor[0], T.nminor[1]);
    const int ngpt = T.ngpt, neta = T.neta, np1 = T.npres + 1;
    uint4 *sitem = smem_q;                                                   // [LCHUNK][nf2][2]
    unsigned *srec = (unsigned *)(smem_q + (size_t)LCHUNK * G.nf2 * 2);    // [LCHUNK][words]
    double *srow = (double *)(srec + (size_t)LCHUNK * G.words);            // [LCHUNK][ROW_DBL]

    // interpolation of the chunk's layers, FP64: the per-layer part ...
    if (tid < nl)
        interp_layer(T, l0 + tid, d, ncol_ld, col_idx, col0, p_lay, p_lev, t_lay, vmr, srec + tid * G.words,
                     srow + tid * ROW_DBL, err);

    // per-lane constants: flavors per regime and the minor-absorber items
    // covering this g-point; NM slots per regime (NM = the table's most items
    // covering one g-point); unused slots point at interval NMN, whose scale
    // is 0, and contributor 0, so every load is unconditional.
    const bool act = tid < ngpt;
    const int g = act ? tid : 0;
    const int gfl0 = T.gpt_flav[2 * g], gfl1 = T.gpt_flav[2 * g + 1];
    unsigned mit[2][NM > 0 ? NM : 1];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int i0 = T.minor_gptr[r][g], n = T.minor_gptr[r][g + 1] - i0;
#pragma unroll
        for (int ii = 0; ii < NM; ++ii) {
            unsigned v = (unsigned)G.nmn << 16;
            if (ii < n) {
                const int iv = T.minor_items[r][2 * (i0 + ii)], ci = T.minor_items[r][2 * (i0 + ii) + 1];
                v = (unsigned)ci | ((unsigned)iv << 16) | ((unsigned)T.minor_par[r][5 * iv + 4] << 24);
            }
            mit[r][ii] = v;
        }
    }
    const unsigned nc0 = (unsigned)T.ncontrib[0], nc1 = (unsigned)T.ncontrib[1];
    const unsigned dp_ = (unsigned)(neta * ngpt), de = (unsigned)ngpt;   // jpress and eta strides of the major table
    __syncthreads();

    // ... then the (layer, item) and (layer, minor interval) parts over the
    // whole block
    for (int k = tid; k < nl * G.nf2; k += blockDim.x) {
        const int j = k / G.nf2;
        interp_item(T, srec + j * G.words, srow + j * ROW_DBL, k - j * G.nf2, sitem + 2 * k);
    }
    for (int k = tid; k < nl * (G.nmn + 1); k += blockDim.x) {
        const int j = k / (G.nmn + 1), m = k - j * (G.nmn + 1);
        if (m < G.nmn) interp_minor(T, G, srec + j * G.words, srow + j * ROW_DBL, m);
        else srec[j * G.words + G.ms + G.nmn] = 0u;   // the scale of unused item slots
    }
    __syncthreads();
    if (!act) return;

    // GU layers at a time: every table load of them (major, Planck fraction
    // or Rayleigh, minor) is issued before any arithmetic; layers past the
    // chunk end repeat its last layer and are not stored
    constexpr int GU = gather_layers<LW>();
    for (int j0 = 0; j0 < nl; j0 += GU) {
        float k[GU][2][4], q[GU][2][4], kr[GU][2][2], km[GU][NM > 0 ? NM : 1][2][2];
        int regv[GU];
#pragma unroll
        for (int u = 0; u < GU; ++u) {
            const int j = min(j0 + u, nl - 1);
            const unsigned *r = srec + j * G.words;
            const int reg = (int)(__builtin_amdgcn_readfirstlane(r[0]) >> 16) - 1;
            regv[u] = reg;
            const int f = reg == 0 ? gfl0 : gfl1;
            const uint4 *Ij = sitem + (size_t)j * G.nf2 * 2;
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                // corners (jpress, eta), (jpress + 1, eta) and their eta + 1 neighbours
                const uint4 a = Ij[(2 * f + t) * 2];
                const unsigned base = a.x + (unsigned)g;
                if (LW) {
                    const float4 v0 = ld4(T.kq, base), v1 = ld4(T.kq, base + dp_);
                    k[u][t][0] = v0.x; k[u][t][2] = v0.y; q[u][t][0] = v0.z; q[u][t][2] = v0.w;
                    k[u][t][1] = v1.x; k[u][t][3] = v1.y; q[u][t][1] = v1.z; q[u][t][3] = v1.w;
                } else {
                    k[u][t][0] = ld1(T.kmaj, base); k[u][t][1] = ld1(T.kmaj, base + dp_);
                    k[u][t][2] = ld1(T.kmaj, base + de); k[u][t][3] = ld1(T.kmaj, base + dp_ + de);
                    const unsigned b0 = __umul24(Ij[(2 * f + t) * 2 + 1].x, (unsigned)ngpt) + (unsigned)g;
                    const float *kt = reg == 0 ? T.kray[0] : T.kray[1];
                    kr[u][t][0] = ld1(kt, b0); kr[u][t][1] = ld1(kt, b0 + de);
                }
            }
            const unsigned nc = reg == 0 ? nc0 : nc1;
#pragma unroll
            for (int ii = 0; ii < NM; ++ii) {
                const unsigned v = reg == 0 ? mit[0][ii] : mit[1][ii];
                const int fv = mi_fv(v);
                const unsigned ci = (unsigned)mi_ci(v);
                // only lanes whose slot holds an item fetch (an empty slot's
                // scale is 0, so its value is never seen): the gathers are
                // bound by the data the texture path returns per lane
                const bool item = mi_iv(v) != G.nmn;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const unsigned b0 = __umul24(Ij[(2 * fv + t) * 2 + 1].x, nc) + ci;
                    float k0 = 0.0f, k1 = 0.0f;
                    if (item) {
                        if (LW) {
                            const float2 m2 = ld2(reg == 0 ? T.km2[0] : T.km2[1], b0);
                            k0 = m2.x;
                            k1 = m2.y;
                        } else {
                            const float *kmt = reg == 0 ? T.kmin[0] : T.kmin[1];
                            k0 = ld1(kmt, b0);
                            k1 = ld1(kmt, b0 + nc);
                        }
                    }
                    km[u][ii][t][0] = k0;
                    km[u][ii][t][1] = k1;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < GU; ++u) {
            const int j = min(j0 + u, nl - 1), l = l0 + j, reg = regv[u];
            const unsigned *r = srec + j * G.words;
            const uint4 *Ij = sitem + (size_t)j * G.nf2 * 2;
            const int f = reg == 0 ? gfl0 : gfl1;
            const float fp1 = __uint_as_float(r[1]), fp0 = 1.0f - fp1;
            float tg = 0.0f, pf = 0.0f, kv = 0.0f;
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const uint4 a = Ij[(2 * f + t) * 2];
                const float fe0 = __uint_as_float(a.y), fe1 = __uint_as_float(a.z);
                const float st = fp0 * (fe0 * k[u][t][0] + fe1 * k[u][t][2]) + fp1 * (fe0 * k[u][t][1] + fe1 * k[u][t][3]);
                tg += __uint_as_float(a.w) * st;
                if (LW) pf += fp0 * (fe0 * q[u][t][0] + fe1 * q[u][t][2]) + fp1 * (fe0 * q[u][t][1] + fe1 * q[u][t][3]);
                else kv += fe0 * kr[u][t][0] + fe1 * kr[u][t][1];
            }
            // minor absorbers of this layer's regime covering g
#pragma unroll
            for (int ii = 0; ii < NM; ++ii) {
                const unsigned v = reg == 0 ? mit[0][ii] : mit[1][ii];
                const int fv = mi_fv(v);
                float kvm = 0.0f;
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const uint4 m = Ij[(2 * fv + t) * 2 + 1];
                    kvm += __uint_as_float(m.y) * km[u][ii][t][0] + __uint_as_float(m.z) * km[u][ii][t][1];
                }
                tg += __uint_as_float(r[G.ms + mi_iv(v)]) * kvm;
            }
            if (j0 + u >= nl) continue;
            if (FB && !row_mask[(size_t)d * nlay + l]) continue;
            const size_t o = ((size_t)d * nlay + l) * ngpt + g;
            tau[o] = tg;
            aux[o] = LW ? pf : kv * __uint_as_float(r[2]);
        }
    }
}

template <bool LW, bool FB>
static void launch_nm(const TablesView &T, int ncols, int nlay, int ncol_ld, const int *col_idx, int col0,
                      const double *p_lay, const double *p_lev, const double *t_lay, const VmrPtrs &vmr, int *err,
                      float *tau, float *aux, hipStream_t st, const unsigned char *row_mask,
                      const unsigned char *col_mask)
{
    const size_t smem = gather_smem(rec_geom(T.nflav, T.nminor[0], T.nminor[1]));
    const unsigned nb = (unsigned)ncols * (unsigned)((nlay + LCHUNK - 1) / LCHUNK);
    const unsigned grid = 8u * ((nb + 7u) / 8u);
#define GO_LAUNCH(NMV)                                                                                        \
    hipLaunchKernelGGL((gas_optics_kernel<LW, NMV, FB>), dim3(grid), dim3(128), smem, st, T, nlay, ncol_ld,  \
                       col_idx, col0, ncols, p_lay, p_lev, t_lay, vmr, err, tau, aux, row_mask, col_mask)
    const int nm = T.maxminor_g;
    if (nm <= 0) GO_LAUNCH(0);
    else if (nm == 1) GO_LAUNCH(1);
    else if (nm == 2) GO_LAUNCH(2);
    else if (nm <= 4) GO_LAUNCH(4);
    else GO_LAUNCH(MAX_MINOR_G);
#undef GO_LAUNCH
}

void launch_gas_optics(const TablesView &T, int ncols, int nlay, int ncol_ld, const int *col_idx, int col0,
                       const double *p_lay, const double *p_lev, const double *t_lay, const VmrPtrs &vmr, int *err,
                       float *tau, float *aux, hipStream_t st, const unsigned char *row_mask,
                       const unsigned char *col_mask)
{
    if (ncols <= 0 || nlay <= 0) return;
    const bool fb = row_mask != nullptr;
    if (T.is_lw) {
        if (fb) launch_nm<true, true>(T, ncols, nlay, ncol_ld, col_idx, col0, p_lay, p_lev, t_lay, vmr, err, tau, aux, st,
                                      row_mask, col_mask);
        else launch_nm<true, false>(T, ncols, nlay, ncol_ld, col_idx, col0, p_lay, p_lev, t_lay, vmr, err, tau, aux, st,
                                    nullptr, nullptr);
    } else {
        launch_nm<false, false>(T, ncols, nlay, ncol_ld, col_idx, col0, p_lay, p_lev, t_lay, vmr, err, tau, aux, st,
                                nullptr, nullptr);
    }
}

}  // namespace ccpprad
