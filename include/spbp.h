/* spbp.h -- C ABI of libspbp, the MI355X (gfx950) Fourier-gridding
 * backprojection behind SPIDER's `BP 3F` / `BP 32F` operations.
 *
 * The reference has no library boundary: the operations BP3F / BP32F
 * (bp3f.f:79, bp32f.f:57) prompt for files and call the accumulation engine
 * BP3FQ / BP32FQ (bp3f.f:298, bp32f.f:433), whose inner kernel is ONELINE
 * (oneline.f:38), and then the finalisation NRMW2 (bp3f.f:1347) + WINDKB2A
 * (bp32f.f:945).  These entry points replace that engine + finalisation seam;
 * the caller (a Fortran `bind(C)` driver or the Python extension) keeps the
 * operation's signature: image series + angle doc in, volume(s) out.
 *
 * Conventions
 *   - Images: nimg x NX x NX, x fastest (= Fortran PROJ(NX,NX)), fp32 or bf16.
 *   - Angles: nimg x 3 doc-file registers (PSI, THETA, PHI) in degrees; the
 *     rotation is CANG(PHI, THETA, PSI) (bp3f.f:1290-1291, cang.f:44).
 *   - Half-grid Fourier volume: complex64 X(0:NX, N, N) and fp32 W(0:NX, N, N),
 *     N = 2 NX, kx fastest, ky/kz in FFT wrap order (Fortran X(0:NX,N2,N2)).
 *   - Output volumes: NX^3 fp32, x fastest.
 *   - `stream` is a hipStream_t (NULL = the legacy default stream).
 *   - Device pointers unless the name ends in _host.
 *   - Return codes mirror the reference's IRTFLG / ERRT numbers where one
 *     exists; 0 = success.  spbp_last_error() gives the message.
 *   - One context per stream; calls on one context are not re-entrant.
 */
#ifndef SPBP_H
#define SPBP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define SPBP_OK             0
#define SPBP_ERR_ALLOC      46   /* ERRT 46: allocation failed (bp32f.f:180-185)            */
#define SPBP_ERR_SIZE      101   /* ERRT 101: ONLY WORKS ON SQUARE IMAGES / bad box size     */
#define SPBP_ERR_ANGLE     102   /* ERRT 102: MISSING ANGLE FOR IMAGE (bp3f.f:1283-1288)     */
#define SPBP_ERR_FFT       201   /* FFT plan / execution failure (fmrs.f:246-258)            */
#define SPBP_ERR_DEVICE    202   /* HIP runtime error                                        */
#define SPBP_ERR_ARG       203   /* invalid argument (NULL context, bad half flag, ...)      */
#define SPBP_ERR_NONFINITE 205   /* NaN / Inf in the images (only with SPBP_CHECK_FINITE)    */

/* image storage types */
#define SPBP_F32  0
#define SPBP_BF16 1

/* option flags */
#define SPBP_TWO_HALVES   0x01  /* BP 32F: two half-set accumulators (bp32f.f:175-178)         */
#define SPBP_EXACT_PI     0x02  /* exact pi in CANG instead of float32(pi)/180 (cang.f:52-53)  */
#define SPBP_TWOSUM_W     0x04  /* compensated (two-sum) accumulation of the weights           */
#define SPBP_CHECK_FINITE 0x08  /* reject images holding NaN / Inf                             */
#define SPBP_TIMING       0x10  /* time each kernel launch with HIP events (spbp_kernel_stats) */
#define SPBP_GEOM_F64     0x20  /* FP64 rotations (from FP64 angles), sample positions, nearest voxels
                                   and table indices: departs from the reference's fp32 geometry
                                   (oneline.f:56-105) to agree with an FP64 computation          */
#define SPBP_HIPFFT       0x80  /* forward 2D transforms by hipFFT on the fully padded images instead
                                   of the library's own two-pass transforms (power-of-two N only;
                                   other sizes always use hipFFT)                              */
#define SPBP_PLAIN_THRESHOLD 0x40 /* NRMW2's W > 0.1 test on the GPU sums alone.  By default the
                                   voxels whose W lies within 2^-11 of 0.1 are decided by
                                   re-summing their weights in the reference's order (see
                                   spbp_threshold_stats)                                       */
#define SPBP_FBS          0x100 /* FBS ON (setmode.f:128,169): RB inserts pre-rotate with RTSF,
                                   Fourier-based bicubic interpolation (rtsf.f, fbs2.f), instead of
                                   RTSQ (bp32f.f:811-822)                                       */

/* kernels reported by spbp_kernel_stats */
#define SPBP_K_INSERT   0   /* ONELINE scatter (units: projections)       */
#define SPBP_K_PAD      1   /* PADD2 (units: projections)                 */
#define SPBP_K_FFT2     2   /* batched 2D forward FFT (units: projections) */
#define SPBP_K_FINALIZE 3   /* one whole finalisation (units: volumes)    */
#define SPBP_K_BIN      4   /* brick work-list build (units: projections) */
#define SPBP_K_GATHER   5   /* central-slice KB gather (units: projections) */
#define SPBP_K_RESOLVE  6   /* NRMW2 threshold re-summation (units: finalisations) */

/* insertion strategies */
#define SPBP_STRAT_AUTO   0   /* the library's choice                                     */
#define SPBP_STRAT_DIRECT 1   /* one thread per sample, FP32 atomics into the half grid   */
#define SPBP_STRAT_BRICK  2   /* LDS-binned bricks: workgroup owns a brick of the grid     */
#define SPBP_STRAT_GATHER 3   /* voxel-driven gather: thread owns a voxel, sums in the
                                 serial reference's order (bit-identical to it, no atomics) */

typedef struct {
    int nx;      /* image edge NX; volume NX^3, padded Fourier grid N = 2 NX          */
    int flags;   /* SPBP_* option flags                                               */
    int device;  /* HIP device ordinal, -1 = current device                           */
    int batch;   /* projections per internal FFT/insert chunk, 0 = automatic          */
    int strategy;    /* SPBP_STRAT_*                                                  */
    int brick_proj;  /* brick strategy: projections per work item, 0 = automatic      */
} spbp_opts;

typedef struct spbp_ctx spbp_ctx;

/* Create a context: allocates and zeroes the accumulators (BP32F's
 * ALLOCATE XE,WE,XO,WO + zeroing, bp32f.f:175-178, 490-500). */
int spbp_create(const spbp_opts* opts, spbp_ctx** out);
int spbp_destroy(spbp_ctx* ctx);
/* Message of the last failure on ctx (ctx == NULL: last failure of this thread). */
const char* spbp_last_error(const spbp_ctx* ctx);
/* Zero the accumulators. */
int spbp_reset(spbp_ctx* ctx, void* stream);

/* Central-slice Kaiser-Bessel gather: the exact adjoint (transpose, complex
 * numbers taken as pairs of reals) of the ONELINE scatter (oneline.f:38-158).
 * For each projection i and lattice sample (I, J), J = -NX+1..NX:
 *   out_bi[i][J mod N][I] = c( sum_direct w X[t] + conj(sum_reflected w X[-t]) )
 * over the taps t, weights w and reflections ONELINE would use for that
 * sample (c = conj for a Friedel-flipped sample); samples ONELINE skips are 0.
 * On an SPBP_GEOM_F64 context the sample positions are FP64, as that
 * context's insertion computes them, so the transpose stays exact.
 * X: caller-owned complex64 half grid, (NX+1) N N (device);
 * out_bi: nimg x N x (NX+1) complex64 (device), ONELINE's BI layout (SPIDER
 * sign and checkerboard applied).  Rotations from angles_deg (host nimg x 3,
 * as spbp_insert) or, when dm is not NULL, dm (host nimg x 9 fp32).
 * No counterpart in the reference: SPIDER's own projector PJ 3G
 * (pj3g.f:125) uses a different kernel; this is the projector whose adjoint
 * is BP 3F's insertion, for gradient and CG reconstruction. */
int spbp_gather(spbp_ctx* ctx, const float* X, int nimg, const double* angles_deg, const float* dm,
                float* out_bi, void* stream);
/* ONELINE scatter of BI-layout spectra (the adjoint of spbp_gather), added
 * into a caller-owned half grid X and, when W is not NULL, weights W.  Same
 * insertion kernels as spbp_insert (BP32FQ's projection loop,
 * bp32f.f:802-913) minus PADD2 and the forward FFT. */
int spbp_scatter(spbp_ctx* ctx, const float* bi, int nimg, const double* angles_deg, const float* dm,
                 float* X, float* W, void* stream);
/* spbp_scatter without a context and without allocating: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * nx and flags (SPBP_EXACT_PI, SPBP_GEOM_F64; others are refused) stand for
 * the context; bi, angles_deg / dm, X and W are as for spbp_scatter, W may be
 * NULL.  Scratch (KB table, geometry, raw spectra, brick lists, queue head,
 * job table) lives in a file-scope device array of 64 MiB that every GPU the
 * library runs on holds until the process ends; boxes whose per-chunk
 * scratch does not fit (NX of 733 and above) are refused with SPBP_ERR_SIZE.  The
 * projections go through in chunks of up to 256, and the call synchronises
 * `stream` after each chunk (its host staging is on the stack).  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU race. */
int spbp_scatter_noalloc(int nx, int flags, const float* bi, int nimg, const double* angles_deg, const float* dm,
                         float* X, float* W, void* stream);
/* The projector applied to a grid: for nimg rotations, the gather of X
 * (spbp_gather's BI spectra), completed to the full C2R input (conjugated,
 * checkerboard (-1)^(I + jp), column 0 by Hermitian symmetry), transformed
 * by hipFFT C2R, scaled by 1 / N^2 and cropped to NX x NX at the padding
 * offset: nimg device images of NX x NX fp32.  No counterpart in the
 * reference (it has no projector matched to ONELINE). */
int spbp_gather_images(spbp_ctx* ctx, const float* X, int nimg, const double* angles_deg, const float* dm,
                       float* imgs, void* stream);
/* The projector's adjoint applied to real images, for power-of-two padded
 * sizes: the scatter (no weights) of the spectra A^T's chain gives for nimg
 * device images of NX x NX fp32 -- the R2C transforms of the zero-padded
 * images (no PADD2 mean) times cx(I) / N^2 (cx = 2 for 0 < I < NX, else 1),
 * column 0 folded as the transpose of its Hermitian completion -- added into
 * X.  Equals spbp_scatter of the BI spectra computed from those images; no
 * counterpart in the reference (it has no projector adjoint). */
int spbp_scatter_images(spbp_ctx* ctx, const float* imgs, int nimg, const double* angles_deg, const float* dm,
                        float* X, void* stream);
/* Insert projections: PADD2 + forward FFT + checkerboard + ONELINE over all
 * rows, per image (the BP32FQ / BP3FQ projection loop, bp32f.f:802-913,
 * bp3f.f:403-493).  half[i] = 0 -> first sample volume (XE), 1 -> second
 * (XO), as RANDLIST does (bp32f.f:544-553); half may be NULL (all in the
 * first).  angles_deg and half are host arrays. */
int spbp_insert(spbp_ctx* ctx, const void* imgs, int dtype, int nimg,
                const double* angles_deg, const int8_t* half, void* stream);
/* Same with precomputed rotations: dm is nimg x 9 fp32 CANG matrices (host);
 * rows 1-2 of each must be orthonormal to 1e-4 (SPBP_ERR_ANGLE otherwise). */
int spbp_insert_dm(spbp_ctx* ctx, const void* imgs, int dtype, int nimg,
                   const float* dm, const int8_t* half, void* stream);
/* Same with host-resident images: uploaded in sub-chunks of about 256 MiB
 * (at most the context's batch) into two device buffers by turns, on a
 * stream of the library's own, so that each upload overlaps the insertion
 * of the sub-chunk before; returns when every image is inserted.  Page-locked
 * images upload by asynchronous DMA; pageable ones are first copied by host
 * threads into two page-locked staging buffers the context keeps (2 x the
 * sub-chunk size), the copy of one sub-chunk overlapping the upload and
 * insertion of the one before.  Device or managed memory is copied as it
 * is. */
int spbp_insert_host(spbp_ctx* ctx, const void* imgs_host, int dtype, int nimg,
                     const double* angles_deg, const int8_t* half, void* stream);

/* Finalise: SYMPLANE0 (bp32f.f:1051) + NRMW2 (bp3f.f:1347: divide by W,
 * inverse 3D FFT) + WINDKB2A (bp32f.f:945) into NX^3 fp32 volumes.
 * out_full = both halves together (BP 3F volume / BP 32F overall volume);
 * out_h1, out_h2 = the two sample volumes (SPBP_TWO_HALVES only).  Any may
 * be NULL.  The accumulators are left unchanged (more inserts may follow). */
int spbp_finalize(spbp_ctx* ctx, float* out_full, float* out_h1, float* out_h2, void* stream);
int spbp_finalize_host(spbp_ctx* ctx, float* out_full, float* out_h1, float* out_h2, void* stream);
/* One volume set to finalise: spbp_finalize's arguments for one context. */
typedef struct {
    spbp_ctx* ctx;
    float* out_full;   /* device NX^3 of ctx's NX, or NULL */
    float* out_h1;
    float* out_h2;
} spbp_final_job;
/* Finalise the volumes of several contexts (distinct, one device, any box
 * sizes, at most 64) as spbp_finalize would, in one pass: the NRMW2
 * re-summation of every volume in shared phases (three host synchronisations
 * per call, four when a band list outgrows its buffer, instead of four per
 * volume), then per wave -- the first requested volume of every context,
 * then the second, then the third -- one normalisation launch over all the
 * wave's volumes, their inverse 3D transforms, and one window, shell-mean and
 * subtraction launch over all of them.  The reference finalises one volume
 * after another (bp32f.f:275-410).  spbp_finalize is this call with one job.
 * With SPBP_TIMING each context records the call's re-summation and
 * finalisation spans.  No counterpart in the reference (one operation, one
 * volume). */
int spbp_finalize_batch(const spbp_final_job* jobs, int njobs, void* stream);

/* Partial volumes = `BP 3F P` / `BP 3F M` (bp3f.f:191-248, 531-739):
 * export writes the post-SYMPLANE0 half-grid (X interleaved complex64,
 * 2 (NX+1) N N floats; W (NX+1) N N floats) of `half` (0, 1, or -1 = sum);
 * import adds (accumulate != 0) or sets such a half-grid into `half`. */
int spbp_export_partial(spbp_ctx* ctx, int half, float* X, float* W, void* stream);
int spbp_import_partial(spbp_ctx* ctx, int half, const float* X, const float* W,
                        int accumulate, void* stream);

/* Raw accumulators of `half` (0 or 1), before SYMPLANE0, same layout.  No
 * counterpart in the reference (its accumulators are internal arrays); for
 * diagnostics and tests. */
int spbp_export_raw(spbp_ctx* ctx, int half, float* X, float* W, void* stream);

/* With SPBP_TIMING: total device milliseconds, launch count and work units
 * of `kernel` (SPBP_K_*) since the last spbp_reset_stats; synchronises on the
 * recorded events.  No counterpart in the reference. */
int spbp_kernel_stats(spbp_ctx* ctx, int kernel, double* total_ms, long* launches, long* units);
int spbp_reset_stats(spbp_ctx* ctx);
/* Measured peaks for the insertion's atomic roofline (no counterpart in the
 * reference), timed on the current device with HIP events on `stream`:
 * SPBP_CAL_LDS_U64: 64-bit LDS atomic adds as the brick kernel issues them
 * (512-thread groups, 64 active lanes, no bank conflicts), in wave-
 * instructions per second; SPBP_CAL_GLOBAL_F32: non-returning fp32 global
 * atomic adds, consecutive lanes on consecutive words, spread over `bytes`
 * of device memory (at least 2 MiB per 256 CUs), in lane operations per
 * second.  Allocates and frees its own buffer; synchronises the stream. */
#define SPBP_CAL_LDS_U64    0
#define SPBP_CAL_GLOBAL_F32 1
int spbp_calibrate(int what, size_t bytes, double* rate, void* stream);

/* NRMW2 (bp3f.f:1355-1366) divides where W > 0.1 and zeroes elsewhere, a
 * test that two summation orders can decide differently when W lies within
 * rounding distance of 0.1.  The library keeps the geometry of every section
 * inserted since the last reset and, when finalising, re-sums the weights of
 * each voxel whose W lies within 2^-11 of 0.1 in the reference's order --
 * images in input order, symmetries, rows J ascending, I ascending, direct
 * before reflected taps (bp32f.f:866-898, oneline.f:56-152), in fp32 (FP64
 * with SPBP_GEOM_F64) -- and takes the decision from that sum.  After
 * spbp_import_partial the history no longer describes the accumulators and
 * the GPU sums decide.  For the last finalised volume: voxels re-summed,
 * decisions that differ from the GPU sums' own, and re-sums that disagreed
 * with the GPU weight beyond rounding (left to the GPU sums; 0 unless the
 * accumulators were changed outside the library).  No counterpart in the
 * reference (its sums are the order). */
int spbp_threshold_stats(const spbp_ctx* ctx, long long* listed, long long* changed, long long* inconsistent);
/* The same re-summation for any n half-grid voxels (index ((kz mod N) N +
 * (ky mod N)) (NX+1) + kx): W as NRMW2 reads it -- SYMPLANE0 pair and, for
 * which = -1, both halves added -- summed in the reference's order, returned
 * as double (the fp32 sum widened, or the FP64 sum with SPBP_GEOM_F64).
 * which: -1 all halves, 0 or 1 one half.  Synchronous.  No counterpart in
 * the reference; for diagnostics and tests. */
int spbp_reference_weights(spbp_ctx* ctx, int which, const long long* voxels, int n, double* w_out, void* stream);

/* Sizes (elements) of one half-grid: complex X count = (NX+1) N N. */
size_t spbp_grid_elems(const spbp_ctx* ctx);

/* RB 3F / RB 32F (bp32f.f:805-837, CALLRTSQ): each image is first rotated by
 * its in-plane angle and shifted (RTSQ with QUADRI interpolation, rtsq.f:56,
 * quadri.f:83) and then inserted as spbp_insert does.  inplane: host
 * nimg x 3 = doc registers 6, 7, 8 (angle in degrees, x shift, y shift). */
int spbp_insert_rb(spbp_ctx* ctx, const void* imgs, int dtype, int nimg, const double* angles_deg,
                   const float* inplane, const int8_t* half, void* stream);
int spbp_insert_rb_host(spbp_ctx* ctx, const void* imgs_host, int dtype, int nimg, const double* angles_deg,
                        const float* inplane, const int8_t* half, void* stream);
/* RTSQ alone (scale 1): nimg NX x NX device images (fp32 / bf16) -> out
 * (device fp32), params host nimg x 3 as above.  Synchronises the stream. */
int spbp_rtsq(const void* imgs, int dtype, int nimg, int nx, const float* params, float* out, void* stream);
/* RTSF alone (FBS ON; rtsf.f:106-205 with scale 1, FBS2 interpolation from
 * the image's Fourier derivatives, fbs2.f:95-245): as spbp_rtsq. */
int spbp_rtsf(const void* imgs, int dtype, int nimg, int nx, const float* params, float* out, void* stream);
int spbp_rtsf_host(const float* imgs_host, int nimg, int nx, const float* params, float* out_host);
/* The same on host memory (fp32 in and out, staged through device memory;
 * synchronous): the transformed images that RB 3F / RB 32F write under the
 * template answer (bp32f.f:517-527, 817-827, WRTVOL of PROJ). */
int spbp_rtsq_host(const float* imgs_host, int nimg, int nx, const float* params, float* out_host);
/* Symmetries (BP 3F / BP 32F "SYMMETRIES DOC", bp32f.f:133-172): nsym doc
 * rows (PSI, THETA, PHI) -> matrices SM by CANG (BUILDS, builds.f:35-63).
 * With nsym > 1 every image is inserted once per DMS = MATMUL(SM_k, DM)
 * (bp32f.f:866-872) and every finalised volume goes through SYMVOL
 * (symvol.f:39, bp32f.f:278-303).  nsym <= 1 switches symmetries off, as a
 * one-row or absent document does in the reference.  One set of symmetries
 * serves a whole operation (insertion and SYMVOL): the call fails with
 * SPBP_ERR_ARG once the accumulators hold sums (after an insert or import,
 * until spbp_reset).  The gather / scatter entry points ignore it. */
int spbp_set_symmetries(spbp_ctx* ctx, int nsym, const double* angles_deg);
/* ---- batching, tuning and graphs (no counterpart in the reference: its
 * projection loop is one OpenMP loop per operation, bp32f.f:802-913) ---- */
/* Plan = insertion strategy (SPBP_STRAT_DIRECT / _BRICK / _GATHER), brick shape
 * (0: 16x16x8, 1: 16x8x8, 2: 8x8x8, 3: 16x12x8, 4: 16x16x10, 5: 16x16x12,
 * 6: 16x16x13 voxels = the LDS bin, every sample evaluated in each bin its
 * taps reach; 7, the default: owner-computes bins, a 16x16x13 LDS region
 * around a 12x12x9 core, each sample evaluated only in the bin whose core
 * holds its nearest voxel) and projections per
 * brick work item (default 128; a launch of fewer than 8 x that many
 * sections halves it, down to 16, so its busiest bricks still split).  spbp_tune times every plan on a sample of
 * up to 1024 of the given images (into scratch grids: the accumulators are
 * untouched) and keeps the fastest; spbp_get_plan reports the plan and, when
 * ms_per_proj is not NULL, the 34 measured ms per projection ([shape][C = 16,
 * 32, 64, 128], then direct, then gather).  The gather plan
 * (SPBP_STRAT_GATHER) is the scatter's transpose: each thread owns a voxel
 * and adds the terms of the chunk's sections in the serial reference's order
 * (projections, symmetries, rows J, columns I, direct before reflected tap,
 * bp32f.f:866-898, oneline.f:53-152), so X and W are the serial reference's
 * fp32 sums bit for bit and identical from run to run; it does several times
 * the arithmetic of the scatter plans. */
int spbp_tune(spbp_ctx* ctx, const void* imgs, int dtype, int nimg, const double* angles_deg, const float* dm,
              void* stream);
int spbp_get_plan(const spbp_ctx* ctx, int* strategy, int* shape, int* brick_proj, float* ms_per_proj);
int spbp_set_plan(spbp_ctx* ctx, int strategy, int shape, int brick_proj);
/* One insertion job: spbp_insert's arguments for one context. */
typedef struct {
    spbp_ctx* ctx;
    const void* imgs;          /* device, nimg x NX x NX of ctx's NX */
    int dtype;                 /* SPBP_F32 / SPBP_BF16 */
    int nimg;
    const double* angles_deg;  /* host nimg x 3, or NULL when dm is given */
    const float* dm;           /* host nimg x 9, or NULL */
    const int8_t* half;        /* host, may be NULL */
} spbp_job;
/* Insert several jobs (distinct contexts, any box sizes, at most 64) as
 * spbp_insert would, with the scatter work of all of them in one persistent
 * kernel launch per round of chunks.  The contexts share one brick shape:
 * the best total when all are tuned, else the first job's. */
int spbp_insert_batch(const spbp_job* jobs, int njobs, void* stream);
/* A batch captured as a HIP graph (pad, FFT, list build, persistent kernel;
 * with reset_first, zeroing of the accumulators first).  The graph keeps its
 * own copy of the geometry; images are read from the jobs' device pointers at
 * every launch.  The graph replays work on its contexts' chunk buffers,
 * brick lists and forward FFT plans, so while it exists every call that
 * would free or replace them fails with SPBP_ERR_ARG: spbp_destroy,
 * spbp_tune, spbp_set_symmetries, spbp_set_plan with another brick shape,
 * and insert / scatter / gather calls that need larger chunk buffers,
 * longer lists or a forward FFT plan for a new chunk size.  Calls that fit
 * the existing buffers may run between launches.  spbp_graph_launch replays
 * it on `stream`; spbp_graph_destroy releases the contexts. */
typedef struct spbp_graph spbp_graph;
int spbp_graph_create(const spbp_job* jobs, int njobs, int reset_first, spbp_graph** out);
int spbp_graph_launch(spbp_graph* g, void* stream);
int spbp_graph_destroy(spbp_graph* g);
/* Host-side tables, exactly as the kernels use them (no GPU needed):
 * the KB interpolation table TABI(0:4999) and FLTB for padded size n2
 * (bp32f.f:193-214, fillbessi0.f:88-117) -- evaluated in REAL arithmetic as
 * the reference does (bit-identical, I1 by Cody's CALCI1 as besi1.f), or with
 * flags & SPBP_GEOM_F64 the FP64 evaluation rounded to fp32 that contexts
 * with FP64 geometry use;
 * the CANG matrices (cang.f:44-82) of n doc-angle triples (PSI, THETA, PHI),
 * n x 9 fp32, with flags & SPBP_EXACT_PI selecting exact pi. */
int spbp_kb_table(int n2, int flags, float* tab, float* fltb);
int spbp_rotations(int n, const double* angles_deg, int flags, float* dm);
/* Make the k-th device allocation of the library from now on (k = 0: the
 * next one) fail as if the device were out of memory; k < 0 cancels.  A
 * failed allocation leaves the context usable: the call returns
 * SPBP_ERR_ALLOC and a later call allocates again.  No counterpart in the
 * reference; for tests of the allocation error path (ERRT 46,
 * bp32f.f:180-185). */
int spbp_debug_fail_alloc(long k);
/* Library version string. */
const char* spbp_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SPBP_H */
