/* holo.h -- C ABI of the MI355X angular-spectrum / Gerchberg-Saxton library (libholo.so).
 *
 * Every entry point replaces one function of ParticleHolography.jl (src/holofunc.jl,
 * src/frequency_filters.jl, src/utils.jl, src/ccl.jl, src/particle_detection.jl) or says that
 * it has no counterpart there; the comment above each declaration names it (file:line). All pointers are device pointers
 * (HIP global memory) unless stated; the caller allocates every input and output. Calls are
 * asynchronous and ordered on the given `stream` (a hipStream_t, NULL = the null stream).
 *
 * Layout: a plane of nx x ny values is stored x-contiguous: element (i, j) at j*nx + i. This
 * is the reference's column-major CuArray{T,2} of size (nx, ny) (Julia dim 1 = x). Complex
 * values are interleaved (re, im) float pairs (ComplexF32). Volumes are nx*ny*slices with
 * the slice index slowest (Julia CuArray{T,3} of size (nx, ny, slices)).
 *
 * Supported FFT lengths (per axis): 2^a, 3*2^a and 5*2^a, up to 4096 (holo_supported_length).
 *
 * Return value: HOLO_OK (0) or an error code; holo_last_error() gives the message of the
 * calling thread's last failure. A handle is not thread safe: use one handle per host thread.
 */
#ifndef HOLO_H
#define HOLO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct holo_context* holo_handle;

enum holo_status {
  HOLO_OK = 0,
  HOLO_ERR_ARG = 1,     /* bad argument / size mismatch (the reference's AssertionError) */
  HOLO_ERR_SIZE = 2,    /* FFT length not supported */
  HOLO_ERR_TYPE = 3,    /* unsupported element type (the reference's ArgumentError) */
  HOLO_ERR_HIP = 4,     /* HIP runtime error */
  HOLO_ERR_NOMEM = 5,   /* device (or, for holo_link_rectangles, host) allocation failed */
  HOLO_ERR_BUDGET = 6   /* request exceeds the configured device-memory budget */
};

/* Element types of volume outputs (return_type in holofunc.jl:158-166). */
enum holo_dtype {
  HOLO_N0F8 = 0, /* FixedPointNumbers N0f8: uint8 bit pattern, value u/255 */
  HOLO_F16 = 1,
  HOLO_F32 = 2,
  HOLO_F64 = 3,
  HOLO_C64 = 4   /* ComplexF32 (complex volume, holofunc.jl:216-232) */
};

/* Precision of a transfer function built by holo_transfer. */
enum holo_tf_precision {
  HOLO_TF_REFERENCE = 0, /* the reference's arithmetic: Float32 sqrt of d, FP64 exp -> CF32 */
  HOLO_TF_FP64_D = 1,    /* FP64 sqrt of the given d values, phase wrapped mod 2pi in FP64 */
  HOLO_TF_FP64 = 2       /* d recomputed in FP64 from (wavlen, dx), phase wrapped in FP64 */
};

/* A transfer function as consumed by the reconstruction / propagation entry points.
 * generated = 1: H(z) is generated on the device in FP64-wrapped form from (z, wavlen, dx)
 *                (no array read); `data` is ignored.
 * generated = 0: `data` is a centred (fftshifted) ComplexF32 nx*ny array, as returned by
 *                holo_transfer / the reference's cu_transfer (types.jl:25-27). */
typedef struct holo_tf {
  const void* data;
  double z;
  double wavlen;
  double dx;
  int generated;
} holo_tf;

/* ---- lifecycle -------------------------------------------------------------------------- */
int holo_create(holo_handle* out, int device);
int holo_destroy(holo_handle h);
const char* holo_last_error(void);
int holo_version(void);
int holo_supported_length(int n);
/* Tuning knobs (0 = library default): slices per batch (round), slices per workgroup of the
 * reconstruction pass that generates H_k (spg_gen; the row pass of projections, the column pass
 * of volumes) and of the pass that forms the output (spg_out), columns per workgroup of column
 * passes (power of two), rows per workgroup of row passes. */
int holo_set_tuning(holo_handle h, int batch, int spg_gen, int spg_out, int cols_per_wg,
                    int rows_per_wg);

/* ---- A2: cu_transfer_sqrt_arr (holofunc.jl:8-15, 30-36) --------------------------------- */
/* d[x,y] = 1 - ((x-nx/2-1)wavlen/nx/dx)^2 - ((y-ny/2-1)wavlen/ny/dx)^2 (1-based x, y), FP64
 * arithmetic in the reference's order, stored as Float32 (dtype HOLO_F32, bit-identical to the
 * reference) or Float64 (HOLO_F64). */
int holo_transfer_sqrt(holo_handle h, void* stream, int nx, int ny, double wavlen, double dx,
                       int dtype, void* out);

/* ---- A3: cu_transfer (holofunc.jl:38-45, 61-67) ----------------------------------------- */
/* H = exp(2 pi i z / wavlen * sqrt(d)) as a centred ComplexF32 array. `d_sqr` (F32 or F64,
 * d_dtype) is read for precisions HOLO_TF_REFERENCE and HOLO_TF_FP64_D; HOLO_TF_FP64 uses
 * (wavlen, dx) only and may pass d_sqr = NULL. */
int holo_transfer(holo_handle h, void* stream, int nx, int ny, double z, double wavlen,
                  const void* d_sqr, int d_dtype, int precision, double dx, void* out);

/* ---- A4: cu_gabor_wavefront (holofunc.jl:80-82) ----------------------------------------- */
int holo_gabor_wavefront(holo_handle h, void* stream, int nx, int ny, const float* holo,
                         void* out);

/* ---- A6: cu_get_reconst_xyprojection (holofunc.jl:248-266) ------------------------------ */
/* proj = min over k = 0..slices-1 of |IFFT2(FFT2(wavefront) .* H_k)|^2 with H_k the transfer
 * to front.z + k*dz.z. With two generated transfers H_k is evaluated per slice in FP64-wrapped
 * form; with two arrays the reference's ComplexF32 running product H0 .* Hdz^k is used. */
int holo_reconst_xyprojection(holo_handle h, void* stream, int nx, int ny,
                              const void* wavefront, const holo_tf* front,
                              const holo_tf* dz, int slices, float* proj);

/* Same, starting from a real intensity hologram: the wavefront sqrt(holo) is fused into the
 * first FFT pass (cu_gabor_wavefront followed by cu_get_reconst_xyprojection). */
int holo_reconst_xyprojection_holo(holo_handle h, void* stream, int nx, int ny,
                                   const float* holo, const holo_tf* front,
                                   const holo_tf* dz, int slices, float* proj);

/* Allocation-free variant of holo_reconst_xyprojection. No counterpart in the reference (whose
 * functions allocate through CUDA.jl's pool); it is for callers that cannot allocate device
 * memory: code under graph capture, tight serving loops, harnesses that forbid hipMalloc. Same
 * arguments, results within FP32 rounding of holo_reconst_xyprojection's. It calls no hipMalloc
 * or hipFree: its workspace is a static 128 MiB device array (one per GPU, held from the
 * library's load until the process ends) whose address holo_create looks up; FFT twiddle
 * tables not cached on the handle are copied into it. Sizes whose spectrum plus one slice
 * (2 planes, 3 for array transfers, of nx*ny*8 bytes) exceed that array return
 * HOLO_ERR_BUDGET; smaller sizes run in rounds that fit it. Every call on one GPU shares the
 * array: calls that overlap on one GPU (other streams, other handles, other threads) race and
 * must be serialized by the caller. */
int holo_reconst_xyprojection_noalloc(holo_handle h, void* stream, int nx, int ny,
                                      const void* wavefront, const holo_tf* front,
                                      const holo_tf* dz, int slices, float* proj);

/* ---- A7 / A8 / A9: cu_get_reconst_vol, cu_get_reconst_complex_vol,
 *      cu_get_reconst_vol_and_xyprojection (holofunc.jl:184-200, 216-232, 300-321) -------- */
/* vol (nx*ny*slices of out_dtype). If proj != NULL (A9, Float32 nx*ny) it receives the min over
 * slices of the stored values (N0f8 as u/255). HOLO_C64 stores psi_k unclamped (A8) and
 * requires proj == NULL. */
int holo_reconst_vol(holo_handle h, void* stream, int nx, int ny, const void* wavefront,
                     const holo_tf* front, const holo_tf* dz, int slices,
                     int out_dtype, void* vol, float* proj);

/* ---- f2: cu_2d_pad and cu_get_reconst_vol_and_xyprojection_padded
 *      (holofunc.jl:139-149, 339-364) ------------------------------------------------------- */
/* out (2nx x 2ny ComplexF32) = mean(in) everywhere, `in` copied at offset (nx/2, ny/2). */
int holo_2d_pad(holo_handle h, void* stream, int nx, int ny, const void* in, void* out);
/* Transfers are 2nx x 2ny; vol is N0f8 nx*ny*slices (the centre window), proj Float32 nx*ny. */
int holo_reconst_vol_and_xyprojection_padded(holo_handle h, void* stream, int nx, int ny,
                                             const void* wavefront, const holo_tf* front,
                                             const holo_tf* dz, int slices,
                                             uint8_t* vol, float* proj);

/* ---- A10: cu_phase_retrieval_holo (holofunc.jl:100-129) --------------------------------- */
/* Gerchberg-Saxton: psi1 = sqrt(I1); `iters` times: psi2 = P_fwd psi1, psi2 <- sqrt(I2) psi2/|psi2|,
 * psi1 = P_bwd psi2, psi1 <- sqrt(I1) psi1/|psi1|. out (ComplexF32 nx*ny) = psi1. With
 * use_graph != 0 the identical iterations run as hipGraphs of up to 16 iterations, captured
 * once per size and tile choice and cached on the handle (they reference only its workspace,
 * into which the amplitudes sqrt(I1) and sqrt(I2) are written, as holofunc.jl:107-108 takes them
 * once per call). */
int holo_phase_retrieval(holo_handle h, void* stream, int nx, int ny, const float* holo1,
                         const float* holo2, const holo_tf* fwd, const holo_tf* bwd,
                         int iters, void* out, int use_graph);

/* ---- A11: cu_asm_prop! (holofunc.jl:383-388) -------------------------------------------- */
/* out = IFFT2(FFT2(in) .* H(tf)); out may alias in. */
int holo_asm_prop(holo_handle h, void* stream, int nx, int ny, void* out, const void* in,
                  const holo_tf* tf);

/* ---- A12: CUFFT fft / ifft (call sites in holofunc.jl) ---------------------------------- */
/* 2-D C2C FFT, forward unnormalized, inverse scaled by 1/(nx*ny) (AbstractFFTs convention).
 * out may alias in. */
int holo_fft2(holo_handle h, void* stream, int nx, int ny, const void* in, void* out,
              int inverse);

/* ---- low-pass filters: cu_rectangle_filter, cu_super_gaussian_filter (frequency_filters.jl:31-74)
 *      and cu_apply_low_pass_filter(!) (frequency_filters.jl:88-107) ---------------------- */
enum holo_lpf_kind {
  HOLO_LPF_RECTANGLE = 0,      /* 1 inside the cutoff circle, 0 outside */
  HOLO_LPF_SUPER_GAUSSIAN = 1  /* exp(-(r/sigma)^6 / 2), one half at the cutoff radius */
};
/* Centred n x n Float32 filter for objects up to prop_dist away (cutoff radius
 * n^2 dx^2 / (wavlen sqrt(4 prop_dist^2 + n^2 dx^2)) samples, Fugal 2009 Eq. 13-15), computed
 * in FP64 with the reference's operation order and centring. */
int holo_lowpass_filter(holo_handle h, void* stream, int kind, double prop_dist, double wavlen,
                        int n, double dx, float* out);
/* out = IFFT2(ifftshift(lpf .* fftshift(FFT2(in)))) for a centred Float32 nx*ny filter;
 * out may alias in (cu_apply_low_pass_filter! passes the same buffer). */
int holo_apply_low_pass_filter(holo_handle h, void* stream, int nx, int ny, void* out,
                               const void* in, const float* lpf);

/* ---- ingest: 8-bit holograms to Float32 (load_gray2float, utils.jl:23-25) with the tutorial's
 *      background removal img .- bkg .+ 64/255 (preprocessings.md:21-23) ------------------- */
/* `count` uint8 images of nx*ny pixels (image m at src + m*nx*ny, raw N0f8 bytes) to Float32:
 * out = Float32((Float64(Float32(v/255)) - bkg) + offset); bkg (Float64 nx*ny, shared by all
 * images) may be NULL, meaning 0, and then offset 0 gives exactly load_gray2float's v/255. */
int holo_ingest_u8(holo_handle h, void* stream, int nx, int ny, int count, const uint8_t* src,
                   const double* bkg, double offset, float* out);

/* ---- background of an image stack: cu_make_background_mode (utils.jl:243-317) and
 *      make_background(...; mode=:mean) (utils.jl:217-224) --------------------------------- */
enum holo_background_kind {
  HOLO_BACKGROUND_MODE = 0,  /* per-pixel most frequent value (smallest on ties) / 255.0 */
  HOLO_BACKGROUND_MEAN = 1   /* per-pixel FP64 sum of v/255 in image order, / count */
};
/* out (Float64 nx*ny) from `count` >= 1 uint8 images (image m at stack + m*nx*ny); the mode
 * takes at most 65535 images per call. */
int holo_make_background(holo_handle h, void* stream, int nx, int ny, int count,
                         const uint8_t* stack, int kind, double* out);

/* ---- connected components and bounding boxes of slices: cu_connected_component_labeling,
 *      cu_find_valid_labels, get_bounding_rectangles (ccl.jl:186-264) and the per-slice part of
 *      particle_bounding_boxes (particle_detection.jl:58-96); cu_dilate (particle_detection.jl:12-42) */
enum holo_ccl_kind {
  HOLO_CCL_F32 = 0,     /* Float32 image: values > 0 are foreground; 8-neighbours with equal
                           values are connected (cu_connected_component_labeling on a Float32 image) */
  HOLO_CCL_BINARY = 1,  /* uint8 / Bool image: nonzero is foreground (a slice of d_bin_vol) */
  HOLO_CCL_N0F8_LE = 2  /* uint8 N0f8 volume thresholded on the fly: foreground where
                           v/255 <= threshold (d_vol .<= threshold) */
};
/* One component of one slice: label = 1 + the smallest offset j*nx + i of its pixels (the
 * reference's label), coordinates 0-based and inclusive (i fast, j slow). */
typedef struct holo_box {
  uint32_t label;
  int32_t i0, j0, i1, j1;
} holo_box;
/* labels (uint32, count slices of nx*ny): 1 + the smallest offset in the pixel's 8-connected
 * component, 0 for background. Images hold at most 2^32-2 pixels. */
int holo_ccl_label(holo_handle h, void* stream, int nx, int ny, int count, const void* img, int kind,
                   double threshold, uint32_t* labels);
/* Boxes of every component of every slice without a label image: slice z's boxes at
 * boxes + z*max_boxes in no particular order, nboxes[z] their number (device int[count]); a
 * slice with more than max_boxes components stores the first max_boxes claimed and reports
 * the full count. */
int holo_ccl_boxes(holo_handle h, void* stream, int nx, int ny, int count, const void* img, int kind,
                   double threshold, int max_boxes, holo_box* boxes, int* nboxes);
/* The same boxes from a label image (get_bounding_rectangles on every valid label). */
int holo_label_boxes(holo_handle h, void* stream, int nx, int ny, int count, const uint32_t* labels,
                     int max_boxes, holo_box* boxes, int* nboxes);
/* out = 3 x 3 in-plane dilation of each slice (uint8 / Bool), false on the slice border. */
int holo_dilate(holo_handle h, void* stream, int nx, int ny, int count, const uint8_t* in, uint8_t* out);
/* Host-only (no device work, no handle): the z-linking of particle_bounding_boxes(_3d)
 * (particle_detection.jl:58-96): gen_particle_neighborhoods on slice 1 (ccl.jl:266-277), one
 * update_particle_neighborhoods! / update_particle_neighborhoods3d! per further slice
 * (ccl.jl:328-409, three_d selects), then finalize_particle_neighborhoods! (ccl.jl:424-466).
 * rects: host int32 (x_min, y_min, x_max, y_max) of every slice's rectangles in order, slice s
 * (1-based) at rects[4*offsets[s-1] .. 4*offsets[s]); offsets has nslices + 1 entries. Every
 * neighbourhood ever started is reported in start order (capacity offsets[nslices] rows):
 * rows (int64 x_min, y_min, z_min, x_max, y_max, z_max, final values), born (its start slice),
 * alive (1 if it survives the unit-cube deletions and finalize); *nrows their number. The
 * reference's dict keys are the caller's: one key per started neighbourhood, in this order. */
int holo_link_rectangles(const int32_t* rects, const int64_t* offsets, int nslices, int three_d,
                         int64_t* rows, int32_t* born, uint8_t* alive, int64_t* nrows);

/* ---- batched reconstruction (no counterpart in the reference: its functions take one
 *      hologram; this runs many cu_get_reconst_xyprojection calls, holofunc.jl:248-266, together) */
typedef struct holo_job {
  int nx, ny;
  const void* src;   /* Float32 hologram (src_real = 1, wavefront sqrt(holo)) or ComplexF32 wavefront */
  int src_real;
  holo_tf front;     /* generated transfers only */
  holo_tf dz;
  int slices;
  float* proj;       /* Float32 nx*ny min-projection */
} holo_job;
/* Reconstructs every job's xy min-projection. Jobs are grouped by (nx, ny, wavlen, dx); one
 * column-pass launch and one row-pass launch per round cover every hologram of a group (up to
 * 32 at a time; workspace of (1 + batch) planes per hologram, chunked to HOLO_BATCH_WS_BYTES).
 * Each result equals the single-hologram call's. */
int holo_reconst_batch(holo_handle h, void* stream, const holo_job* jobs, int njobs);
/* Batched Gerchberg-Saxton: many cu_phase_retrieval_holo calls (holofunc.jl:100-129) together.
 * Pairs are grouped by size; each pass of an iteration is one launch over up to 32 pairs.
 * Pairs may have different transfers and iteration counts. Each result equals the single
 * call's. */
typedef struct holo_gs_job {
  int nx, ny;
  const float* holo1;
  const float* holo2;
  holo_tf fwd;       /* holo1 -> holo2 */
  holo_tf bwd;       /* holo2 -> holo1 */
  int iters;
  void* out;         /* ComplexF32 nx*ny */
} holo_gs_job;
int holo_phase_retrieval_batch(holo_handle h, void* stream, const holo_gs_job* jobs, int njobs);
/* Batched GS runs the iterations that are identical for a chunk of pairs (all but each pair's
 * last) as hipGraphs of up to 16 iterations, captured once per size, chunk layout and workspace
 * and cached on the handle (the amplitudes sqrt(I1), sqrt(I2) are written into the workspace, so
 * the graphs reference no caller memory); enable = 0 launches every pass directly. A chunk holds
 * as many pairs as keep its fields and amplitudes (16 B/px) within 256 MiB, at most 32
 * (HOLO_GS_CHUNK sets the count). No counterpart in the reference. */
int holo_set_gs_graphs(holo_handle h, int enable);
/* Caps the workspace the batched calls plan for (holograms are processed in chunks that fit;
 * a single hologram larger than the cap still runs). 0 = HOLO_BATCH_WS_BYTES or 8 GiB. */
int holo_set_workspace_limit(holo_handle h, size_t bytes);
/* The device workspace the handle holds now (bytes): it grows on demand to the largest call's
 * need (spectrum, a round of column-transformed slices, GS tables) and is freed by
 * holo_destroy. No counterpart in the reference (CUDA.jl's pool owns its temporaries). */
int holo_workspace_size(holo_handle h, size_t* bytes);
/* Grows the handle's workspace to at least `bytes` now (no counterpart in the reference). A
 * call that needs more than the handle holds grows it itself, which synchronizes the device;
 * a server reserves its slots' workspaces up front so that no call does that mid-stream. */
int holo_reserve_workspace(holo_handle h, size_t bytes);
/* Per-size tile choices (0 = default): slices per batch, slices per column / row workgroup,
 * columns per workgroup, rows per workgroup. The global holo_set_tuning overrides them. The
 * size's kernel variants (chosen by holo_autotune or shipped: elements per thread, one or two
 * exchange buffers in the column pass, rows loaded one slice ahead in the row pass) stay. */
int holo_set_tuning_for_size(holo_handle h, int nx, int ny, int batch, int spg_gen, int spg_out,
                             int cols_per_wg, int rows_per_wg);
/* The effective choices for a size (defaults resolved) as {batch, spg_gen, spg_out, cols, rows}. */
int holo_get_tuning_for_size(holo_handle h, int nx, int ny, int slices, int* out5);
/* A size's kernel variants (no counterpart in the reference): pass A's exchanges (1 two buffers,
 * -1 one, 0 two where they leave as many workgroups resident per CU; for sides whose columns are
 * 1024 tall, 2 and 3: the last exchange by cross-lane swaps and one LDS exchange per slice in
 * buffers alternating by slice, with the twiddles in LDS (2) or in registers (3)) and pass B's row loads
 * (1 one slice ahead, 2 one slice ahead with the twiddles in registers, 3 as 2 with every row
 * exchange through LDS instead of the last one by cross-lane swaps, 0 or -1 in place).
 * set != 0 stores *col_db and *row_pf for the size; set == 0 reads the size's values. */
int holo_size_variants(holo_handle h, int nx, int ny, int* col_db, int* row_pf, int set);
/* Times candidate tile configurations of both reconstruction passes on the device (HIP events,
 * synchronous) and keeps the fastest for this size; the tiles in force before the call stay
 * unless the winner is at least 2% faster. out5 (may be NULL) receives the tiles in force
 * afterwards, defaults resolved as by holo_get_tuning_for_size with slices = 0. */
int holo_autotune(holo_handle h, void* stream, int nx, int ny, double wavlen, double dx, int* out5);

/* ---- measurement (no counterpart in the reference, which has no timers) ---------------- */
/* Passes timed with HIP events when profiling is on: the fused reconstruction's column pass
 * (pass A) and row pass (pass B), the forward spectrum, the GS propagation column pass and
 * the GS row pass. Units: slices for the reconstruction passes, launches otherwise. Events
 * are recorded on the caller's stream around each launch (not inside graph captures). */
enum holo_pass {
  HOLO_PASS_RECON_COLS = 0,
  HOLO_PASS_RECON_ROWS = 1,
  HOLO_PASS_FWD_FFT = 2,
  HOLO_PASS_PROP_COLS = 3,
  HOLO_PASS_GS_ROWS = 4
};
int holo_set_profiling(holo_handle h, int enable);
/* Waits for the recorded events; returns the accumulated time (ms), launches and units of a
 * pass since the last reset. */
int holo_profile_read(holo_handle h, int pass, double* total_ms, long long* launches,
                      long long* units, int reset);

#ifdef __cplusplus
}
#endif
#endif /* HOLO_H */
