// Device-side data structures and helpers shared by the gfx950 kernels.
//
// Layout in HBM (all device tables FP32 unless stated):
//   LW, eta pairs (eta + 1 past the last eta row holds 0, never read):
//   kq                : [ntemp][npres+1][neta][ngpt] float4 {k(eta), k(eta+1),
//                       Planck fraction(eta), (eta+1)}: the two eta corners of
//                       both tables in one 16-byte load, g-point innermost
//   km2[r]            : [ntemp][neta][ncontrib_r] float2, contributor innermost
//   SW, plain:
//   kmaj              : [ntemp][npres+1][neta][ngpt]
//   kmin[r]           : [ntemp][neta][ncontrib_r]
//   kray[r]           : [ntemp][neta][ngpt] Rayleigh
//   totplnk           : [nband][nplanck]
//   vmr_ref           : [ntemp][ngas+1][2]  FP64 (feeds the FP64 interpolation)
// Per-call work buffers, g-point innermost so a wavefront's lanes (one lane
// per g-point of one column) touch one contiguous row:
//   tau / pfrac / tau_ray : [col][lay][gpt]
//   mask                  : [col][lay][MASK_WORDS] u32 bits, bit g of word g/32
//   cloud band optics     : [lay][col][band] (LW: absorption depth; SW: tau, co-albedo, g)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ccpprad {

constexpr int MAX_GAS = 6;
constexpr int MAX_FLAV = 16;
constexpr int MAX_MINOR = 32;      // minor intervals per regime
constexpr int MAX_MINOR_G = 8;     // minor intervals covering one g-point, per regime
constexpr int MAX_BAND = 32;
constexpr int MAX_GPT = 128;       // g-points per column handled by one block
constexpr int MASK_WORDS = 4;      // 128 bits of McICA mask per (col, lay)
constexpr int MAX_ANG = 4;
constexpr int MAX_LAY = 1024;

// rte-rrtmgp physical constants (mo_gas_optics_constants)
constexpr double AVOGAD = 6.02214076e23;
constexpr double M_DRY = 0.028964;
constexpr double M_H2O = 0.018016;
constexpr double GRAV = 9.80665;
constexpr double CP_DRY = 1004.64;

// Read-only view of the device tables, passed by value as a kernel argument.
struct TablesView {
    int is_lw, ngas, nband, ngpt, ntemp, npres, neta, nflav, nplanck;
    double press_ref_log1, press_ref_log_delta, press_ref_trop_log;
    double temp_ref_min, temp_ref_delta, totplnk_delta;
    double press_min, press_max, temp_min, temp_max;
    const double *temp_ref;    // [ntemp] FP64
    const float4 *kq;          // LW major absorption and Planck fraction, eta pairs
    const float2 *km2[2];      // LW minor absorption per regime, eta pairs
    const float *kmaj;         // SW major absorption
    const float *kmin[2];      // SW minor absorption per regime
    const float *kray[2];      // SW Rayleigh per regime
    const float *totplnk;
    const float *opt_fit;      // [nband][2]
    const double *vmr_ref;
    const double *flav_ratio;  // [ntemp][nflav][2] vmr_ref(gas1)/vmr_ref(gas2) per regime
    const int *flavor;         // [nflav][2]
    const int *gpt_flav;       // [ngpt][2] 0-based flavor per regime
    const int *g2b;            // [ngpt] band of g-point (0-based)
    const int *band_lims;      // [nband][2] 0-based first, last
    int nminor[2];
    int ncontrib[2];
    int maxminor_g;            // most minor intervals covering one g-point in either regime
    const int *minor_par[2];   // [n][5]: idx_minor, density, idx_scaling, complement, flavor(0-based)
    const int *minor_gptr[2];  // [ngpt+1] CSR offsets into minor_items
    const int *minor_items[2]; // [..][2]: interval, contributor index (0-based)
    const float *toa_norm;     // SW [ngpt] toa source scaled to solcon (FP32), written per call
    const double *toa_norm64;  // SW [ngpt] FP64 copy
};

struct LutView {
    int nband, nsize_liq, nsize_ice;
    float liq_lwr, liq_step, ice_lwr, ice_step, liq_upr, ice_upr;
    float liq_istep, ice_istep;               // 1 / step, formed in FP64
    const float *extliq, *ssaliq, *asyliq;   // [nband][nsize_liq]
    const float *extice, *ssaice, *asyice;   // [nband][nsize_ice] (roughness selected)
    const float *coaliq, *coaice;            // co-albedo 1 - ssa, formed in FP64 at init
    int pade;                                // Pade form: coefficients in place of the tables
    const float *pcoef;                      // [phase][nband][3 regimes][16]: extinction 0-5,
                                             // co-albedo 6-10, asymmetry 11-15
    float pbnd[2][3][4];                     // [phase][extinction, co-albedo, asymmetry] regime bounds
};

__device__ __forceinline__ int d_min(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int d_max(int a, int b) { return a > b ? a : b; }

// interpolate1D of the band Planck table (FP32), table row for band b.
__device__ __forceinline__ float planck_band(const TablesView &t, int b, float T)
{
    float v0 = (T - (float)t.temp_ref_min) / (float)t.totplnk_delta;
    int iv = (int)v0;
    float frac = v0 - (float)iv;
    int idx = d_min(t.nplanck - 1, d_max(1, iv + 1));
    const float *tab = t.totplnk + (size_t)b * t.nplanck;
    float a0 = tab[idx - 1], a1 = tab[idx];
    return a0 + frac * (a1 - a0);
}

// base[idx] for a wavefront-uniform base and a 32-bit element index, with the
// byte offset formed in 32 bits: the load can then take the base in SGPRs
// and the offset in one VGPR (no 64-bit address add per load).  The caller
// guarantees idx * sizeof(T) < 2^32.
template <typename T>
__device__ __forceinline__ T ldu(const T *base, unsigned idx)
{
    return *(const T *)((const char *)base + idx * (unsigned)sizeof(T));
}

// 1/x from v_rcp_f32 (about 1 ulp); x must be a normal number.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// e^{-x} from v_exp_f32 (x >= 0).
__device__ __forceinline__ float exp_neg(float x) { return __builtin_amdgcn_exp2f(-1.4426950408889634f * x); }

// e = e^{-x} and E = 1 - e^{-x} for x >= 0, E with small relative error: its
// Taylor series (7 terms, truncation below 2e-9 relative) under x = 0.25,
// 1 - e^{-x} (no cancellation) above.
__device__ __forceinline__ void exp_pair(float x, float &e, float &E)
{
    e = exp_neg(x);
    const float s = x * (1.0f - x * (0.5f - x * (1.0f / 6.0f - x * (1.0f / 24.0f - x * (1.0f / 120.0f -
                    x * (1.0f / 720.0f - x * (1.0f / 5040.0f)))))));
    E = x < 0.25f ? s : 1.0f - e;
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false),
                            __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// FP64 sum over the 64 lanes of a wavefront, returned to every lane (all
// lanes must be active).  Within each row of 16 lanes: quad swaps (lane^1,
// lane^2), then the half-row and row mirrors; the four row sums are then
// read out and added in a fixed order, so every lane gets the same value.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);  // row_half_mirror
    v += dpp_f64<0x140>(v);  // row_mirror
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// Sum of lane pairs by the gfx950 half-wave / row-pair swaps: lanes whose
// lane bit (bit 5 for H32, bit 4 otherwise) is 0 get a(L) + a(L ^ m), the
// others b(L ^ m) + b(L).  v_permlane32_swap exchanges the upper 32 lanes of
// its first operand with the lower 32 of its second; v_permlane16_swap does
// the same for the odd rows of the first and the even rows of the second.
template <bool H32>
__device__ __forceinline__ double swap_sum(double a, double b)
{
    const unsigned ah = (unsigned)__double2hiint(a), al = (unsigned)__double2loint(a);
    const unsigned bh = (unsigned)__double2hiint(b), bl = (unsigned)__double2loint(b);
    const auto h = H32 ? __builtin_amdgcn_permlane32_swap(ah, bh, false, false)
                       : __builtin_amdgcn_permlane16_swap(ah, bh, false, false);
    const auto l = H32 ? __builtin_amdgcn_permlane32_swap(al, bl, false, false)
                       : __builtin_amdgcn_permlane16_swap(al, bl, false, false);
    return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
}

// Quad-level pair sum: lanes whose bit is 0 get a(L) + a(partner), the
// others b(L) + b(partner); CTRL is the quad_perm swapping that bit.
template <int CTRL>
__device__ __forceinline__ double xsel_sum(double a, double b, bool bit)
{
    const double keep = bit ? b : a, send = bit ? a : b;
    return keep + dpp_f64<CTRL>(send);
}

// FP64 sums over the 64 lanes of N = 8 or 16 per-lane FP32 values at once
// (all lanes active).  Each exchange step halves the number of values a lane
// carries, so the N sums cost about what two single sums do.  Lane L gets
// the sum of value
//   N = 16: 8 bit1 + 4 bit0 + 2 bit4 + bit5,   N = 8: 4 bit0 + 2 bit4 + bit5
// of L (lanes differing only in the other bits hold the same, bit-identical
// sum); wave_sum_idx gives that index.
template <int N>
__device__ __forceinline__ double wave_sum_n(const float (&v)[N])
{
    static_assert(N == 8 || N == 16, "8 or 16 values");
    const int lane = __lane_id();
    double a[N / 2];
#pragma unroll
    for (int i = 0; i < N / 2; ++i) a[i] = swap_sum<true>((double)v[2 * i], (double)v[2 * i + 1]);
    double b[N / 4];
#pragma unroll
    for (int i = 0; i < N / 4; ++i) b[i] = swap_sum<false>(a[2 * i], a[2 * i + 1]);
    double d;
    if (N == 16) {
        const double c0 = xsel_sum<0xB1>(b[0], b[1], lane & 1), c1 = xsel_sum<0xB1>(b[2], b[3], lane & 1);
        d = xsel_sum<0x4E>(c0, c1, (lane >> 1) & 1);
    } else {
        d = xsel_sum<0xB1>(b[0], b[1], lane & 1);
        d += dpp_f64<0x4E>(d);    // quad_perm [2,3,0,1]: lanes L, L ^ 2
    }
    d += dpp_f64<0x128>(d);       // row_ror:8: lanes L, L ^ 8
    d += dpp_f64<0x124>(d);       // row_ror:4: lanes L, L + 4, L + 8, L + 12 of the row
    return d;
}

template <int N>
__device__ __forceinline__ int wave_sum_idx(int lane)
{
    return N == 16 ? 8 * ((lane >> 1) & 1) + 4 * (lane & 1) + 2 * ((lane >> 4) & 1) + ((lane >> 5) & 1)
                   : 4 * (lane & 1) + 2 * ((lane >> 4) & 1) + ((lane >> 5) & 1);
}

// the lanes holding each of the N sums once (bits 2-3 zero, bit 1 too for N = 8)
template <int N>
__device__ __forceinline__ bool wave_sum_writer(int lane)
{
    return (lane & (N == 16 ? 12 : 14)) == 0;
}

}  // namespace ccpprad
