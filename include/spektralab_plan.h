/* spektralab_plan.h -- the render plan: every constant and table one render of the chain reads.
 *
 * The engine's float64 set-up (film set-up, print set-up, per-run constants, CAM16 tables) is
 * narrowed once into this POD struct.  The HIP kernels are fed from its float32 members; the
 * float64 members carry the same quantities before narrowing, for the float64 CPU oracle and the
 * precision study.  spkx_session_plan() (spektralab_ext.h) copies a session's plan out so a
 * checker can run the same chain on the CPU.
 *
 * Reference counterparts of the members:
 *   film side      src/film/film_internal.hpp:81-100 (FilmSetup::Impl), kernels/film_shared.h:52-90
 *   halation       src/fx/fx_internal.hpp:47-55 (HalationPlan), src/fx/halation.cpp:104-195
 *   grain          src/fx/fx_internal.hpp:61-71 (GrainConstants), kernels/fx_shared.h:73-88
 *   print side     src/print/internal.hpp:172-182 (PrintRun), kernels/print_shared.h:29-112
 *   scanner        src/fx/scanner.hpp:14-18
 * Layout: little-endian LP64, natural alignment; SPK_PLAN_VERSION changes with any layout change.
 */
#ifndef SPEKTRALAB_PLAN_H
#define SPEKTRALAB_PLAN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPK_PLAN_MAGIC 0x504B5053u /* "SPKP" */
#define SPK_PLAN_VERSION 6u

#define SPK_NL 81        /* wavelength samples 380..780 nm, 5 nm */
#define SPK_NL3 243      /* [lambda][channel] */
#define SPK_PLAN_MAX_L 256
#define SPK_PLAN_MAX_K 1024
#define SPK_PLAN_MAX_MIX 8
#define SPK_CMAX_ROWS 64
#define SPK_CMAX_COLS 720
#define SPK_CMAX_N (SPK_CMAX_ROWS * SPK_CMAX_COLS)
#define SPK_PLAN_MAX_EDR 256  /* EDR tone-map nodes */

/* Grain draw modes (kernels/fx_shared.h:70-72). */
#define SPK_GRAIN_SUM 0u
#define SPK_GRAIN_PLANES 1u
#define SPK_GRAIN_SIMPLE 2u

/* One term of a Gaussian mixture M(x) = sum_k w_k G_{sigma_k}(x), per channel, in pixels. */
typedef struct spk_mix_component {
    double weight[3];
    double sigma[3];
} spk_mix_component;

typedef struct spk_plan {
    uint32_t magic, version;
    uint32_t width, height;
    double pitch_um;       /* pixel pitch of this tier, micrometres */
    double rho_tier;       /* tier long edge / full long edge */
    double fir_crossover;  /* per channel: 0 < sigma < fir_crossover -> FIR, else recursive (3.0: reference) */

    /* ---- film side: upsample, exposure compensation, halation, log, curves, grain ---- */
    uint32_t L;            /* spectral table size */
    uint32_t K;            /* film curve points */
    float a[9];            /* input matrix A (RGB -> XYZ adapted), row-major */
    double a64[9];
    float ec_gain;         /* f32(2^exposure_compensation_ev); multiply skipped when 1 */
    uint32_t ae_active;    /* auto exposure: the tier image is multiplied by ae_gain before the upsample */
    float ae_gain;         /* f32(2^ae_ev) */
    uint32_t pad0;
    double ae_ev;          /* the metered EV of the session's auto_exposure_method */

    /* input transfer decode of the tier image (first stage; mode: 0 sRGB, 1 ProPhoto, 2 Adobe, 3 BT.709) */
    uint32_t decode_active, decode_mode;
    float bt709_bp, bt709_exp;
    /* geometry: the tier image (src_w x src_h) resampled to width x height (crop, rotation, quarter
     * turns, flips; bilinear at float64 source coordinates) */
    uint32_t geometry_active, src_w, src_h, quarter_turns, flip_h, flip_v;
    double crop[4];        /* x, y, w, h as fractions of the source */
    double rotation_deg;
    /* highlight boost (boost_ev != 0) and back-reflection headroom, both on the exposure after EC,
     * each from the frame maximum at that point */
    uint32_t boost_active, headroom_active;
    double boost_ev;
    float headroom_ev;
    uint32_t pad3;
    double lens_blur_px;   /* camera lens blur sigma (lens_blur_um / pitch_um), 0 = none */

    uint32_t scatter_active, n_scatter;
    spk_mix_component scatter[4];
    double scatter_a[3], scatter_b[3];     /* raw <- a raw + b M(raw) */
    uint32_t bounce_active, n_bounce;
    spk_mix_component bounce[SPK_PLAN_MAX_MIX];
    double mix_a[3], mix_b[3];             /* raw <- a raw + b H(raw) */

    /* DIR couplers: chi := M^T s(D), s = D (or dmax - D on a positive); chi := the DIR diffusion
     * mixture of chi; D := F(L - chi; y0) */
    uint32_t couplers_active, coupler_positive;
    float coupler_m[9];                    /* M[c][m], row-major, donor c, receiver m */
    float coupler_dmax[3];
    uint32_t n_dir;
    spk_mix_component dir[4];              /* core + exponential-surrogate tail */

    uint32_t grain_active;  /* draws happen */
    uint32_t grain_mode;    /* SPK_GRAIN_* */
    uint32_t grain_positive, grain_shift, grain_seed, grain_k, grain_n_sub;
    uint32_t grain_dye;     /* dye clouds: PLANES draw, plane l blurred with grain_dye_sigma[l*3 + c], summed */
    float grain_inv_n_sub;
    float dmin_l[9], sat_d[9], count[9], uniformity[3], dmin_c[3];  /* [l*3 + c]; SIMPLE uses [c] */
    double grain_final_blur;  /* sigma px, 0 = none */
    double grain_amount;
    /* micro-structure: G *= phi, phi := lognormal(mean 1, std grain_micro_sigma) per pixel and
     * channel c on stream 100 + c with the film seed, blurred with grain_micro_blur px when that
     * exceeds 0.4; with dye clouds or micro-structure the D_min shift follows: G += -dmin_c */
    uint32_t grain_micro;
    float grain_micro_sigma;
    double grain_micro_blur;
    double grain_dye_sigma[9];  /* [l*3 + c] px, 0 = that plane passes unblurred */

    /* ---- print side: enlarger, gain, paper, scan, levels, to_rgb, gamut, unsharp, encode ---- */
    uint32_t printed;       /* !scan_film */
    uint32_t paper_k;
    uint32_t spectral_fp16; /* 81-bin print and scan integrals as FP16 MFMA products, FP32 accumulation */
    uint32_t pad2;
    /* Per-layer maps of the curve arguments, identity by default (a build-defined superset of the
     * reference's single density_curve_gamma, for the differentiable render):
     * film D_c = F_c(film_gamma_c L + film_offset_c); paper D_c = P_c(paper_gamma_c l + paper_offset_c). */
    double film_gamma[3], film_offset[3], paper_gamma[3], paper_offset[3];
    float enl_eps[SPK_NL3], enl_beta[SPK_NL], enl_wt[SPK_NL3], enl_g[3], enl_o[3];
    double enl_eps64[SPK_NL3], enl_beta64[SPK_NL], enl_wt64[SPK_NL3], enl_g64[3], enl_o64[3];
    float k_gain;
    uint32_t levels_active;
    double k_gain64;
    float levels_m, levels_q;
    float scan_eps[SPK_NL3], scan_beta[SPK_NL], scan_wt[SPK_NL3];
    double scan_eps64[SPK_NL3], scan_beta64[SPK_NL], scan_wt64[SPK_NL3];
    float to_rgb[9];
    double to_rgb64[9];

    /* glare (after the scan, before to_rgb): F := lognormal(mean, std) per pixel on stream 200 with
     * glare_seed; F := G_sigma(F); XYZ_c += (F / 100) tint_c */
    uint32_t glare_active, glare_seed;
    float glare_mean, glare_std;
    double glare_sigma;
    float glare_tint[3];
    uint32_t pad1;

    /* CAM16-UCS gamut compression (kernels/print_shared.h:69-95) */
    uint32_t gamut_active;
    uint32_t lc_active;
    float m_to_xyz[9], m_to_rgb[9], m16[9], m16_inv[9], d_rgb[3];
    float f_l, n_bb, n_cb, n_, z, a_w, c, n_c, j_first, j_last, h_first, h_step, knee[3], lc[3], j_white;
    double cam64[32]; /* float64 of the scalars above, same order: f_l n_bb n_cb n z a_w c n_c j_first j_last
                         h_first h_step knee0..2 lc0..2 j_white d_rgb0..2 */
    double m_to_xyz64[9], m_to_rgb64[9], m16_64[9], m16_inv64[9];

    double scanner_lens_blur_px;   /* plain pixels */
    double unsharp_sigma;          /* 0.7 rho */
    double unsharp_amount;         /* 0.7 */
    uint32_t cctf_active, cctf_mode;
    float m_nid[9];
    double m_nid64[9];

    /* scanner levels (after the scan integral): Y' := clamp(m Y + q, 0, 1); XYZ *= Y' / (Y + 1e-10) */
    /* (levels_active, levels_m, levels_q above) */
    /* EDR tone map on luminance after the unsharp mask: y := max(r_y . rgb, 0); outside [toe, shoulder]
     * y' := 2^(interpolated target_log2_y at (log2 y - lo) / span); rgb *= y' / (y + 1e-10) */
    uint32_t edr_active, edr_n;
    float edr_r_y[3], edr_domain_lo, edr_domain_span, edr_toe_y, edr_shoulder_y;
    float edr_table[SPK_PLAN_MAX_EDR];

    /* ---- tables ---- */
    float table[SPK_PLAN_MAX_L * SPK_PLAN_MAX_L * 3];      /* L x L x 3 [i][j][c] */
    double table64[SPK_PLAN_MAX_L * SPK_PLAN_MAX_L * 3];
    float curves[8 * SPK_PLAN_MAX_K];                      /* [x K][inv K][y 3K][y0 3K], y channel-major */
    double curves_x64[SPK_PLAN_MAX_K], curves_y64[3 * SPK_PLAN_MAX_K], curves_y064[3 * SPK_PLAN_MAX_K];
    float grain_axis[3 * SPK_PLAN_MAX_K], grain_inv[3 * SPK_PLAN_MAX_K], grain_layers[9 * SPK_PLAN_MAX_K];
    float paper[5 * SPK_PLAN_MAX_K];                       /* x[K], r[K], y[K][3] */
    double paper_x64[SPK_PLAN_MAX_K], paper_y64[3 * SPK_PLAN_MAX_K];
    float cmax[SPK_CMAX_N];
    double cmax64[SPK_CMAX_N];
} spk_plan;

#ifdef __cplusplus
}
#endif

#endif /* SPEKTRALAB_PLAN_H */
