// Host-side declarations of the HIP kernels' launchers and their by-value argument structs.
//
// Memory layout on the device (one frame of w x h):
//   rgb input     interleaved float32 x 3 (the ABI's layout), 12 B/px
//   film planes   planar float32: channel c of pixel i at p[c * w * h + i]  (exposure, mixture
//                 accumulators, grain, the cached negative)
//   print planes  planar float32 (the scanned RGB before and after the unsharp blur)
//   output        RGBA16, row pitch `stride` pixels, 8 B/px
// Planar planes make every line of every channel a contiguous (rows) or w-strided (columns)
// float sequence, which is what the blur passes stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spektralab_plan.h"

namespace spk::k {

inline constexpr int kFirMaxR = 64;                 // sigma < 21.5 at truncate 3
inline constexpr int kFirMaxTaps = 2 * kFirMaxR + 1;

// One channel of a FIR pass (taps -R..R).
struct FirArgs {
    uint32_t w, h, r;
    uint32_t acc_mode;   // 0 NONE: out = v; 1 ZERO: out = 0 + wt v; 2 PLANE: out = acc + wt v
    float weight;
    float taps[kFirMaxTaps];
};

// One channel of a Young-van Vliet pass along lines of length n, processed in nch chunks of m.
// The carries between chunks use powers of the companion matrix A (row-major 3x3, state
// (w_i, w_{i-1}, w_{i-2})): pw[d] = A^(m 2^d) (d < 8: up to 256 chunks), last = A^(n - (nch-1) m).
struct IirArgs {
    uint32_t w, h;       // plane size
    uint32_t n, m, nch;  // line length, chunk length, chunk count
    uint32_t acc_mode;   // column pass only (rows: NONE)
    float weight;
    uint32_t adjoint;    // 1: the transposed pass (zero starts, edge terms moved to the far ends)
    double B, B1, B2, B3;
    double pw[8][9];
    double last[9];
};

struct UpsampleArgs {
    uint32_t n, L;
    float a[9];
    float ec_gain;
    uint32_t ec_active;
    float pre_gain;           // auto exposure: the input is multiplied by pre_gain first
    uint32_t pre_active;
    uint32_t decode;          // input decode of each value first: 0 none, 1 + mode (mode 0 sRGB, 1 ProPhoto,
                              // 2 Adobe, 3 BT.709)
    float bt709_bp, bt709_exp;
};

// Geometry resample (interleaved 3-channel): output (ox, oy) samples the source at
// (cx + col[2 ox] + row[2 oy], cy + col[2 ox + 1] + row[2 oy + 1]), float64, bilinear, clamped.
struct GeometryArgs {
    uint32_t src_w, src_h, dst_w, dst_h;
    double cx, cy;
    uint32_t decode;          // the taps decoded first (as UpsampleArgs)
    float bt709_bp, bt709_exp;
};

// Search hints of a non-decreasing table axis: kHintBins equal bins over [first, last) of the axis;
// entry b is the number of axis values <= the bin's lower edge, the point from which a short
// walk finds upper_bound(axis, v) for a v in that bin.
inline constexpr uint32_t kHintBins = 1024;

// The fused develop kernel: halation mix, log10, density curves, grain draw.
struct DevelopArgs {
    uint32_t w, h, K;
    uint32_t mix_active;      // raw := mix_a raw + mix_b acc
    float mix_a[3], mix_b[3];
    uint32_t grain_active;    // grain draws (grain_mode: SPK_GRAIN_SUM, _PLANES or _SIMPLE)
    uint32_t grain_mode;
    uint32_t grain_shift;     // SUM: subtract D_min,c from the sum
    float inv_n_sub;          // SIMPLE: G := draw (1 / N_sub) - D_min,c
    float* planes;            // PLANES: draw value of sub-layer l, channel c, pixel i at [(3 l + c) n + i]
    uint32_t write_density;   // also write the pre-grain density (grain_amount != 1)
    uint32_t seed, positive, gk;
    float dmin_l[9], sat_d[9], count[9], uniformity[3], dmin_c[3];
    float d_per_n[9];         // sat_d / count in binary32 (set by set_grain_scale)
    // DIR couplers: stage 1 writes L to out and chi := M^T s(D) to density; stage 2 reads L from H
    // and the blurred chi from acc and develops D := F(L - chi; y0) (then grain as usual).
    uint32_t coupler_stage;   // 0: no couplers
    uint32_t coupler_positive;
    float coupler_m[9], coupler_dmax[3];
    float fg[3], fo[3];       // curve argument map fg L + fo (identity: 1, 0)
    // Axis 0: the curves' x; axis 1 + c: the grain axis of channel c.  Bin of v:
    // (v - hint_lo[a]) hint_scale[a]; hint_scale[a] = 0 (or hints null): bisect instead.
    const uint16_t* hints;    // 4 x kHintBins
    float hint_lo[4], hint_scale[4];
    float* counts;            // grain: the Poisson counts drawn, 9 planes: draw 3 c + l of pixel i at
                              // [(3 c + l) n + i] (or null)
    float* lambdas;           // grain: the Poisson means of those draws, same layout (or null)
};

// The per-draw density of one grain of sub-layer l, channel c: D_lc / n_lc rounded once in
// binary32, as the draw's scale (D / n) sat needs it.
inline void set_grain_scale(DevelopArgs& a) {
    for (int i = 0; i < 9; ++i) a.d_per_n[i] = a.count[i] > 0.0f ? a.sat_d[i] / a.count[i] : 0.0f;
}

// Per-run constants of the enlarger (they follow the live print fields), passed by value.
// eps and beta (and the scan's) are pre-scaled by -log2(10): transmittance = exp2(eps . d + beta).
struct EnlargerArgs {
    float eps[SPK_NL3], beta[SPK_NL], wt[SPK_NL3], g[3], o[3];
    float k_gain;
    float glare_tint[3];      // XYZ += (F / 100) tint with the glare field F when it is given
    float pg[3], po[3];       // paper argument map pg l + po (identity: 1, 0)
    uint32_t levels_active;   // scanner levels on the scanned XYZ: Y' := clamp(m Y + q, 0, 1)
    float levels_m, levels_q;
};

// Pipeline constants of the print side, device-resident (uploaded with the pipeline's tables).
struct PrintConsts {
    uint32_t paper_k;
    uint32_t printed;         // 0: scanned film, the scan integral reads the negative itself
    float scan_eps[SPK_NL3], scan_beta[SPK_NL], scan_wt[SPK_NL3];
    float to_rgb[9];
    float m_to_xyz[9], m_to_rgb[9], m16[9], m16_inv[9], d_rgb[3];
    float f_l, n_bb, n_cb, n_, z, a_w, c, n_c, j_first, j_last, h_first, h_step, knee[3], lc[3], j_white;
    float cz, inv_cz, f_l_qrt, chroma_k, hundred_over_f_l;  // derived on the host in float64
    uint32_t lc_active, gamut_active, rows, cols;
    uint32_t cctf_active, cctf_mode;
    float m_nid[9];
    float unsharp_a;          // (1 + a), the unsharp combine's first weight; second is -a
    float unsharp_b;
    uint32_t unsharp_active;
    uint32_t edr_active, edr_n;  // EDR tone map before the output matrix (table of edr_n log2 targets)
    float edr_r_y[3], edr_domain_lo, edr_domain_span, edr_toe_y, edr_shoulder_y;
    float edr_table[SPK_PLAN_MAX_EDR];
    // search hints of the paper curve axis (see kHintBins); paper_hint_scale 0: bisect instead
    float paper_hint_lo, paper_hint_scale;
    uint16_t paper_hints[kHintBins];
};

// One 81-bin spectral integral R_c = sum_l exp2(e(l) . d + beta(l)) wt_c(l) with the products and
// the sum on the matrix cores: per wavefront of 64 pixels, 11 steps of v_mfma_f32_16x16x32_f16
// over 8 bins each (bins 81..87 are zero).  Lane l = 16g + j owns pixel l and supplies its own
// transmittances as the B operand (B[k = 8g + s][col j] = T(pixel l, bin 8m + s)); the A operand
// is wt block-diagonal in g (A[row 4g + c][k in group g] = wt_c, zero elsewhere), so the result
// C[4g + c][j], held by lane 16g + j, is R_c of that lane's own pixel.  Lane l holds
// A[row l%16][k = 8(l/16) + s]: channel row (l%16)%4 of wt when (l%16)/4 == l/16, else the zero
// row 3.
// wt is scaled per channel by 2^-e_c (max |wt_c| in [0.5, 1)) and stored as IEEE binary16 bits;
// the exp2 arguments carry +shift (beta_sh = f32(beta) + shift) so that fp16 keeps relative
// precision down to T = 2^-(24 + shift); unscale_c = 2^(e_c - shift).  Densities are clamped to
// [lo, hi], which keeps every exp2 result below 2^15.5 (finite in fp16: a zero of A times an
// infinite B would otherwise spread NaN to other pixels of the wavefront).
struct SpecMfma {
    float beta_sh[88];
    uint16_t wt[4][88];
    float unscale[3];
    float lo, hi;
    uint32_t pad[3];
};

// ---- batched (many frames per launch) forms of the pointwise stages -----------------------------
// One job per frame; a launch covers every frame's pixels in 256-pixel tiles: tile_prefix[f] is
// the first tile of frame f (tile_prefix[nframes] = total), and persistent workgroups stride over
// the tile list.  seeds (device, may be null) overrides each frame's seed: develop reads
// seeds[2 f] (the film seed), the glare draw seeds[2 f + 1] (the print seed).
struct DevelopJob {
    const float* H;
    const float* acc;
    float* out;
    float* density;
    const float *curves, *gaxis, *ginv, *glayers, *lgam;
    DevelopArgs a;
};

struct PrintJob {
    const float* neg;
    float* rgb;
    const PrintConsts* pc;
    const float *paper, *cmax, *glare;
    uint32_t w, h;
    EnlargerArgs e;
};

inline constexpr uint32_t kBatchTile = 256;

// Launchers (stream-ordered, no synchronisation). Planes as documented above.
void upsample(const float* rgb, float* H, const float* table, const UpsampleArgs& a, hipStream_t s);
// Input decode of n values (the meter image).
void decode(const float* in, float* out, uint32_t n, uint32_t mode, float bp, float ex, hipStream_t s);
void geometry(const float* src, float* dst, const double* col, const double* row, const GeometryArgs& a,
              hipStream_t s);
// *out := the largest of n floats (NaN ignored; -inf when none); partial: kMaxBlocks floats.
inline constexpr uint32_t kMaxBlocks = 1024;
void plane_max(const float* in, size_t n, float* partial, float* out, hipStream_t s);
void boost(float* p, size_t n, float x0, float r, float a, float k, hipStream_t s);
// planar 3-channel planes of n pixels
void headroom(float* p, uint32_t n, float clip, float ev, hipStream_t s);
// The passes run on nplanes contiguous planes of a.w x a.h with the same arguments.
void fir_v(const float* in, float* out, const FirArgs& a, uint32_t nplanes, hipStream_t s);
void fir_h(const float* in, float* out, const float* acc, const FirArgs& a, uint32_t nplanes, hipStream_t s);
// Two FIR components of a mixture in one pass: out = (a + weight[0] G_{r0}(in)) + weight[1] G_{r1}(in)
// with a = 0 (first_mode 1, ZERO) or acc (first_mode 2, PLANE); 1 <= r[j] <= kFirPairMaxR, each
// component with its own taps.  Bit-identical to fir_v + fir_h of each component in turn.
struct FirPairArgs {
    uint32_t w, h;
    uint32_t r[2];
    uint32_t first_mode;
    float weight[2];
    float taps[2][kFirMaxTaps];
};
inline constexpr int kFirPairMaxR = kFirMaxR;
inline constexpr int kFirPairTileW = 256, kFirPairTileH = 16;  // output tile of one workgroup
// Floats per LDS row of the pair kernel's tile for radii up to R: at least the padded staged
// width, and 16 mod 64 so that the four rows a wavefront reads fall in distinct banks.
__host__ __device__ constexpr int pair_stride(int R) {
    const int t = kFirPairTileW + 2 * R, s = t + (t >> 4);
    return s + ((16 - s) & 63);
}
void fir_pair(const float* in, float* out, const float* acc, const FirPairArgs& a, uint32_t nplanes, hipStream_t s);
// The reverse pass of two consecutive bounce components with one sigma each for the three
// channels (planar 3-channel planes; FIR blurs are symmetric, so G_j(g_acc) is the transpose):
// g_H += w_{0c} G_0(g_acc) + w_{1c} G_1(g_acc) in place, and part[j] row t, column c := the sum
// over output tile t of plane c of G_j(g_acc) H; rows fir_pair_tiles(w, h) .. nrows - 1 are
// zeroed (the caller keeps fir_pair_tiles(w, h) <= nrows).  a.first_mode and a.weight are unused.
struct FirPairAdj {
    const float* H;
    double* part[2];
    uint32_t nrows;
    float w[2][3];
};
uint32_t fir_pair_tiles(uint32_t w, uint32_t h);
void fir_pair_adjoint(const float* g_acc, float* g_H, const FirPairArgs& a, const FirPairAdj& adj, hipStream_t s);
// Recursive passes: a.m is the compile-time chunk length iir_row_chunk(line length) /
// iir_col_chunk(column height) returns (0: too long for one workgroup), a.nch = ceil(a.n / a.m)
// (the column pass takes up to 64 chunks, rows up to 256).  iir_row filters `lines` contiguous
// lines of a.n samples.
// Both passes in one kernel for 1 <= a.r <= kFir2dMaxR (a.acc_mode applies to the result).
inline constexpr int kFir2dMaxR = 8;
void fir2d(const float* in, float* out, const float* acc, const FirArgs& a, uint32_t nplanes, hipStream_t s);
int iir_row_chunk(uint32_t n);
int iir_col_chunk(uint32_t h);
void iir_row(const float* in, float* out, const IirArgs& a, uint32_t lines, hipStream_t s);
void iir_col(const float* in, float* out, const float* acc, const IirArgs& a, uint32_t nplanes, hipStream_t s);
// Transposes a w x h single-channel plane into h x w; with acc_mode, the transposed value v is
// combined as out = (mode ZERO ? 0 : acc) + weight v.
void transpose(const float* in, float* out, const float* acc, uint32_t w, uint32_t h, uint32_t acc_mode, float weight,
               hipStream_t s);
// out_c = a_c x_c + b_c y_c on n-element single-channel planes (y == nullptr: y reads as 0).
void lincomb(const float* x, const float* y, float* out, uint32_t n, float a, float b, hipStream_t s);
void copy(const float* x, float* out, size_t n, hipStream_t s);
void develop(const float* H, const float* acc, float* out, float* density, const float* curves, const float* gaxis,
             const float* ginv, const float* glayers, const float* lgam, const DevelopArgs& a, hipStream_t s);
// Grain micro-structure field: 3 planes of lognormal(1, sigma) on streams 100 + c (seed_ptr, device,
// may be null, overrides seed); then G := G phi + shift_c over 3 planes of n (phi null: G + shift_c).
void micro_field(float* phi, uint32_t n, float sigma, uint32_t seed, const uint32_t* seed_ptr, hipStream_t s);
void grain_finish(float* G, const float* phi, uint32_t n, const float shift[3], hipStream_t s);
// glare: the blurred glare field (one plane) or nullptr.  rx (6 planes, may be nullptr): the
// enlarger response R and the scanned XYZ before clamping, per channel, for the reverse pass.
void print_pointwise(const float* neg, float* rgb, const PrintConsts* pc, const EnlargerArgs& ea, const float* paper,
                     const float* cmax, const float* glare, uint32_t w, uint32_t h, hipStream_t s, float* rx = nullptr);
// The same with both spectral integrals on the matrix cores (enl: enlarger, scan: scan operands).
void print_pointwise_fp16(const float* neg, float* rgb, const PrintConsts* pc, const EnlargerArgs& ea,
                          const SpecMfma* enl, const SpecMfma* scan, const float* paper, const float* cmax,
                          const float* glare, uint32_t w, uint32_t h, hipStream_t s);
// One lognormal glare value per pixel (stream 200); seed_ptr (device, may be null) overrides seed.
void glare_field(float* out, uint32_t n, float mean, float sd, uint32_t seed, hipStream_t s,
                 const uint32_t* seed_ptr = nullptr);
// Batched develop / print over nframes jobs (device arrays), `grid` persistent workgroups of
// `block` threads (block divides kBatchTile).
void develop_batch(const DevelopJob* jobs, const uint32_t* tile_prefix, uint32_t nframes, const uint32_t* seeds,
                   uint32_t grid, uint32_t block, hipStream_t s);
void print_batch(const PrintJob* jobs, const uint32_t* tile_prefix, uint32_t nframes, uint32_t grid, uint32_t block,
                 hipStream_t s);
// VALU issue-rate probes (probe.hip): blocks x 256 threads, each running valu_probe_chains()
// independent chains of iters instructions of one kind; out: one float per thread.
inline constexpr int kProbeFmaF32 = 0, kProbeExpF32 = 1, kProbeFmaF64 = 2, kProbeMadU64 = 3;
uint32_t valu_probe_chains();
void valu_probe(int kind, uint32_t iters, uint32_t blocks, float* out, hipStream_t s);

void philox_test(const uint32_t* ctr, const uint32_t* key, uint32_t* out, uint32_t n, hipStream_t s);
void poisson_test(const float* lam, float* out, uint32_t n, uint32_t stream, uint32_t seed, const float* lgam,
                  hipStream_t s);
// The unsharp blur (FIR radius 1..kFir2dMaxR, taps in f) fused with the combine and encode.
void unsharp_encode2d(const float* rgb, uint16_t* out, uint32_t stride, const PrintConsts* pc, const FirArgs& f,
                      uint32_t w, uint32_t h, hipStream_t s);
void unsharp_encode(const float* rgb, const float* blurred, uint16_t* out, uint32_t stride, const PrintConsts* pc,
                    uint32_t w, uint32_t h, hipStream_t s);

// ---- the adjoint of the differentiable render (adjoint.hip) -------------------------------------
// Gradient slots of the plan quantities (float64, device): the quantity layout of the
// differentiable render.
inline constexpr uint32_t kAdjEcGain = 0;        // exposure-compensation gain
inline constexpr uint32_t kAdjMixA = 1;          // bounce mix a_c (raw := a raw + b acc)
inline constexpr uint32_t kAdjMixB = 4;          // bounce mix b_c
inline constexpr uint32_t kAdjBounceW = 7;       // bounce component weights [k][c], k < 3
inline constexpr uint32_t kAdjBounceS = 16;      // bounce component sigmas [k][c] (pixels)
inline constexpr uint32_t kAdjFilmGamma = 25;    // film curve argument map: gamma_c, offset_c
inline constexpr uint32_t kAdjFilmOffset = 28;
inline constexpr uint32_t kAdjKGain = 31;        // print gain
inline constexpr uint32_t kAdjPaperGamma = 32;   // paper curve argument map: gamma_c, offset_c
inline constexpr uint32_t kAdjPaperOffset = 35;
inline constexpr uint32_t kAdjUnsharp = 38;      // unsharp amount a: (1 + a) x - a G(x)
inline constexpr uint32_t kAdjReserved = 39;     // (zero)
inline constexpr uint32_t kAdjWt = 40;           // enlarger weights wt[3 l + c], 243 of them
inline constexpr uint32_t kAdjQuantities = kAdjWt + SPK_NL3;

// The reverse-pass kernels write per-workgroup partial sums of their quantities into `part`
// (device float64): a kernel with NQ quantities over n pixels writes adj_blocks(n) rows of NQ
// (part[b * NQ + j]); adj_wt writes adj_wt_blocks(n) rows of 243.  adj_reduce then sums, for every
// quantity q, the column terms[q] names, in a fixed order: the gradients are deterministic.
inline constexpr uint32_t kAdjBlock = 256;      // threads per workgroup of the pointwise kernels
inline constexpr uint32_t kAdjWtBlocks = 1024;  // workgroups of the enlarger-weight reduction
struct AdjTerm {
    uint64_t offset;   // first partial of the column in `part`
    uint32_t nblocks;  // rows (0: the quantity's gradient is 0)
    uint32_t stride;   // doubles between rows
};
uint32_t adj_blocks(uint32_t n);
uint32_t adj_wt_blocks(uint32_t n);
// The sigma gradient of a FIR-tile component (radius 1..kFir2dMaxR) by central differences, fused:
// part[b * 3 + c] := sum over tile b of plane p (channel c = c0 + p) of
// scale_c g (G_plus in - G_minus in), for b < fir2d_tiles(w, h); rows up to nrows zeroed.
uint32_t fir2d_tiles(uint32_t w, uint32_t h);
// The same for a separable FIR component (blur.hip): fir_v2 writes the vertical passes of both tap
// sets; fir_h_dot runs both horizontal passes and writes part[b * 3 + c] for its
// fir_h_dot_blocks(w, h, nplanes) workgroups (channel c = plane), zeroing rows up to nrows.
uint32_t fir_h_dot_blocks(uint32_t w, uint32_t h, uint32_t nplanes);
void fir_v2(const float* in, float* outp, float* outm, const FirArgs& plus, const FirArgs& minus, uint32_t nplanes,
            hipStream_t s);
void fir_h_dot(const float* inp, const float* inm, const float* g, const FirArgs& plus, const FirArgs& minus,
               const float scale[3], uint32_t nplanes, uint32_t nrows, double* part, hipStream_t s);
// Returns hipErrorInvalidValue, launching nothing, for a radius outside 1..kFir2dMaxR.
hipError_t adj_fir2d_dot(const float* in, const float* g, const FirArgs& plus, const FirArgs& minus,
                         const float scale[3], uint32_t nplanes, uint32_t c0, uint32_t nrows, double* part,
                         hipStream_t s);

// enc (interleaved) := transfer(m_nid (ua lin + ub blur)) (blur null: lin), pre-quantisation.
void adj_encode_fwd(const float* lin, const float* blur, float* enc, const PrintConsts* pc, float ua, float ub,
                    uint32_t n, hipStream_t s);
// g_v (3 planes) from g_enc (interleaved).  Partials (1): d/d unsharp amount.
void adj_encode_bwd(const float* lin, const float* blur, const float* g_enc, float* g_v, const PrintConsts* pc, float ua,
                    float ub, uint32_t n, double* part, hipStream_t s);
// g_neg, g_R (3 planes each) from g_lin, at the forward values rx (print_pointwise's 6 planes).
// Partials (7): k_gain, paper gamma 0-2, paper offset 0-2.
void adj_print_bwd(const float* neg, const float* rx, const float* g_lin, const float* glare, float* g_neg, float* g_R,
                   const PrintConsts* pc, const EnlargerArgs& ea, const float* paper, const float* cmax, uint32_t n,
                   double* part, hipStream_t s);
// Partials (243): sum_px g_R_c T_l at [3 l + c].
void adj_wt(const float* neg, const float* g_R, const EnlargerArgs& ea, uint32_t n, double* part, hipStream_t s);
// g_H, g_acc (acc null: g_acc unused) from the gradient of the grain sum (or of the density when
// grain is off).  Partials (12): mix a 0-2, mix b 0-2, film gamma 0-2, film offset 0-2.
void adj_develop_bwd(const float* H, const float* acc, const float* counts, const float* g_G, float* g_H, float* g_acc,
                     const float* curves, const float* gaxis, const float* ginv, const float* glayers,
                     const DevelopArgs& a, uint32_t n, double* part, hipStream_t s);
// g_rgb (interleaved) from g_H.  Partials (1): d/d EC gain.
void adj_upsample_bwd(const float* rgb, const float* g_H, float* g_rgb, const float* table, const UpsampleArgs& a,
                      double* part, hipStream_t s);
// One bounce component (3 planes each): g_H_c += w_c t_c.  Partials (3): <t_c, H_c>.
void adj_bounce(const float* t, const float* H, float* gH, const float w[3], uint32_t n, double* part, hipStream_t s);
// Partials (3): scale_c <g_c, fp_c - fm_c>.
void adj_dot_diff(const float* g, const float* fp, const float* fm, const float scale[3], uint32_t n, double* part,
                  hipStream_t s);
// gq[q] := the reduction of terms[q] (device table of nq entries), q < nq.
void adj_reduce(const AdjTerm* terms, uint32_t nq, const double* part, double* gq, hipStream_t s);

// The print-preview table lookup (session.hip) on the planar negative into RGBA16.
struct LutArgs {
    uint32_t w, h, size;
    float lo[3], s[3];  // s = f32(1 / (hi - lo))
};
void print_lut(const float* neg, const float* lut, uint16_t* out, uint32_t stride, const LutArgs& a, hipStream_t s);

// The static scratch region (scratch.hip): kStaticScratchBytes of device memory per GPU, held from
// the code object's load to the end of the process.  static_scratch_init(device) looks up the
// region's address on that device once (hipGetSymbolAddress with the device current);
// static_scratch(device) returns it afterwards (null before the look-up).
inline constexpr size_t kStaticScratchBytes = size_t(512) << 20;
hipError_t static_scratch_init(int device);
void* static_scratch(int device);

// Tier downscale passes on interleaved 3-channel images (session.hip).
void ds_blur_v(const float* src, float* dst, const float* taps, uint32_t w, uint32_t h, int r, hipStream_t s);
void ds_blur_h(const float* src, float* dst, const float* taps, uint32_t w, uint32_t h, int r, hipStream_t s);
void ds_resample(const float* src, float* dst, const uint32_t* xi, const float* xt, const uint32_t* yi, const float* yt,
                 uint32_t in_w, uint32_t out_w, uint32_t out_h, hipStream_t s);

}  // namespace spk::k
