// Host/device shared declarations of the fastfib_amd engine (not part of the
// public C-ABI in include/fastfib_amd.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "ff_arith.h"

// Items per launch passed by value in the kernel arguments (no descriptor
// uploads); bigger batches are cut into launches of this many.
#define FF_MAXB 4

// One transform array for the outer passes. The forward top pass reads
// `src` (an operand's limbs as 32-bit words, nw of them) and writes `data`
// ([3][L] words, prime-major); later passes work on `data` in place.
struct FFXform {
    const uint32_t *src;
    uint32_t nw;
    uint32_t *data;
};
struct FFXformSet {
    FFXform x[FF_MAXB];
};

// One product for the inner kernels: fa holds a's forward transform (or, for
// one-block transforms, receives it from asrc) and gets the product's; fb /
// bsrc = nullptr squares.
// pair != 0 (multi-pass plans only): fa, fb hold the transforms of a and b;
// fa gets u = cu0 X + cu1 Y + cu2 Z and, with pair bit 2, fb gets v =
// cv0 X + cv1 Y + cv2 Z, combined pointwise before the inverse transform.
// X, Y, Z = a^2, ab, b^2, or with pair bit 4 (degree 4) x^2, xy, y^2 for
// x = a^2 + 2ab, y = a^2 + b^2. pair = 1 (one output, degree 2, weights 0..2)
// runs in the general product's middle round; the others in k_inner_pair
// (weights below 8).
// bspec (one-block plans): 1 when fb holds b's spectrum, tile position i at
// fb[prime L + i], computed once for every product sharing b; 2 for the item
// computing it: a's forward transform (from asrc) stored that way to fa.
struct FFConv {
    const uint32_t *asrc, *bsrc;
    uint32_t anw, bnw;
    uint32_t *fa;
    const uint32_t *fb;
    int pair;
    uint8_t cu[3], cv[3];
    int bspec;
};
struct FFConvSet {
    FFConv c[FF_MAXB];
};

// Parameters of one outer (strided, four-step) pass over blocks of Lb words.
struct FFPass {
    uint32_t L;         // full transform length (words per prime)
    uint32_t Lb;        // length of one block this pass transforms
    uint32_t R;         // rows = Lb / Sb
    uint32_t Sb;        // column stride inside the block
    uint32_t nblocks;   // L / Lb
    uint32_t tpb;       // column tiles per block = Sb / C
    uint32_t rootstep;  // FF_NBASE / Lb
    uint32_t rstep3;    // FF_NBASE / R (radix-3 layer twiddles)
    int rho;            // R = 2^rho or 3 * 2^rho
    int r3;
    int C, logC, ld;
    int top;            // forward: read limbs; inverse: apply the CRT scale
    int rl;             // levels per round (2..4)
    int bw;             // coefficient width (forward top pass)
    uint32_t sc[3];     // R^2 / L mod p (inverse top pass)
};

// Parameters of the inner kernels: contiguous blocks of S = 2^s words, or
// the whole transform in one block when small = 1 (L <= FF_TW, may be 3*2^k).
struct FFInner {
    uint32_t L;
    uint32_t S;
    int s;
    int small;
    int r3;         // small only
    uint32_t rstep3;
    int bw;         // small only: coefficient width
    uint32_t sc[3]; // small only: CRT scale
    int rl;         // levels per round of the pair kernel (k_inner_pair)
};

// A product's full schedule.
struct FFPlan {
    uint32_t L;
    int bw;
    int r3;
    int lg;          // L = 2^lg or 3 * 2^lg
    uint64_t conv;   // product coefficients
    int npass;       // outer passes (0: one small kernel)
    FFPass pass[3];
    FFInner inner;
};

// One product's CRT: residues [3][L] -> nout limbs, one thread per limb in
// blocks of 256 limbs (ff_crt.hip).
struct FFCrtJob {
    const uint32_t *y;       // residues [3][L]
    uint64_t *out;           // nout limbs
    uint32_t *lb;            // look-back workspace: 2 + one word per block of 256 limbs, zeroed
    unsigned long long *lenkey;  // (top nonzero limb + 1) << 8 | its bit length
    uint64_t nout;
    uint64_t conv;
    uint32_t L;
    int bw;
};
struct FFCrtSet {
    FFCrtJob j[FF_MAXB];
};

// Work-list items: one per transform array (outer passes) or product (inner
// kernels, CRT), each with its own schedule; a flat grid is cut among them
// by a start array (item k owns blocks [start[k], start[k+1])).
struct FFOuterItem {
    FFPass ps;
    FFXform x;
};
struct FFInnerItem {
    FFInner in;
    FFConv c;
};

// Kernel kinds for the live per-launch timer (ff_prof_*).
enum FFKernelKind {
    FK_FWD_TOP, FK_FWD, FK_INNER, FK_SMALL, FK_INV, FK_INV_TOP,
    FK_CRT_PACK, FK_CRT_SCAN, FK_CRT_APPLY, FK_STEP, FK_CARRY, FK_NORM, FK_LINEAR, FK_DEC_LEAF, FK_EARLY, FK_STEP_CRT,
    FK_SCHOOL, FK_QUEUE, FK_COUNT
};

// A batch as one persistent launch over a task table (k_queue, ff_ntt.hip;
// the table is built by ff_batch.hip). A task is one workgroup's share of
// one pass of one item, exactly a block of the corresponding list launch.
// Tasks are listed in an order in which every dependency points to an
// earlier task; workgroups claim them in that order with one atomic, wait
// for their dependencies (counters of finished tasks per item), run the
// body and count themselves finished. Outer kinds: FQ_OUTER | inv | top << 1
// | r3 << 2.
enum FFQueueKind {
    FQ_OUTER = 0, FQ_INNER = 8, FQ_INNER_SMALL = 9, FQ_INNER_SMALL_R3 = 10, FQ_CRT = 11, FQ_SCHOOL = 12,
    FQ_SCHOOL_WAVE = 13, FQ_FILL = 14
};
#define FQ_NONE 0xffffffffu
struct FFQTask {
    uint32_t kind;
    uint32_t off;     // byte offset of the item in the batch's descriptor buffer (an FFOuterItem, FFInnerItem,
                      // FFCrtJob, FFSchoolItem, or the output pointer of a zero product)
    uint32_t sub;     // block within the item, prime-major; wave groups: products in the group (1 to 4)
    uint32_t done;    // counter this task increments when it ends (FQ_NONE: none)
    uint32_t dep[2];  // wait until counter dep[k] reaches need[k] (FQ_NONE: no dependency)
    uint32_t need[2];
};
// ctr: FQ_HEADER words (the claim counter of part q at FQ_LINE * q, the
// count of finished workgroups at FQ_LINE * FQ_MAXQ), then counter c at
// FQ_HEADER + c; zero before the first launch, and zeroed again by the
// launch's last workgroup.
#define FQ_MAXQ 8
#define FQ_LINE 32  // words: one 128-byte line per claim counter
#define FQ_HEADER (FQ_LINE * (FQ_MAXQ + 1))
hipError_t ff_launch_queue(const FFQTask *tasks, uint32_t ntasks, const char *desc, uint32_t *ctr, uint32_t nctr,
                           size_t lds_bytes, double bytes, hipStream_t st);
// Dynamic LDS the queue kernel needs for a CRT / schoolbook / one-wave schoolbook task.
size_t ff_queue_lds_crt(void);
size_t ff_queue_lds_school(void);
size_t ff_queue_lds_school_wave(void);

// Schoolbook products below the NTT threshold (ff_school.hip): both
// operands shorter than FF_SCHOOL_LIMBS limbs, one workgroup per product.
#define FF_SCHOOL_LIMBS 200
struct FFSchoolItem {
    const uint64_t *a, *b;   // b == a: a square
    uint32_t na, nb;         // normalized lengths, both < FF_SCHOOL_LIMBS
    uint64_t *out;           // nout = na + nb + 1 limbs written
    unsigned long long *key; // (top nonzero limb + 1) << 8 | its bit length; 0 for zero
};
struct FFSchoolSet {
    FFSchoolItem j[FF_MAXB];
};
extern "C" hipError_t ff_launch_school_list(const FFSchoolItem *items, int n, double bytes, hipStream_t st);
extern "C" hipError_t ff_launch_school_set(const FFSchoolSet *set, int n, double bytes, hipStream_t st);
// Items of at most 128 output limbs (na + nb + 1 <= 128), one per wavefront.
#define FF_SCHOOL_WAVE_OUT 128
extern "C" hipError_t ff_launch_school_wave_list(const FFSchoolItem *items, int n, double bytes, hipStream_t st);

extern "C" {
// ff_engine.hip: when enabled, brackets each launch with HIP events on its
// stream and books its algorithmic bytes under its kind.
void ff_prof_begin(int kind, hipStream_t st);
void ff_prof_end(int kind, double bytes, hipStream_t st);
// the algorithmic 32-bit multiplies of the launch the next ff_prof_end closes
void ff_prof_mults(double mults);

// ff_ntt.hip
int ff_tables_init(void);
const FFTables *ff_tables_dev(void);
void ff_prime_consts(FFPrime out[3]);
uint32_t ff_inv_scale(uint32_t L, int i);  // R^2 / L mod p_i: the CRT-ready inverse scale
int ff_plan_build(FFPlan *pl, uint32_t L, int bw, int lg, int r3, uint64_t conv, int S_log, int tile_words);
hipError_t ff_launch_forward(const FFPlan *pl, const FFXformSet *xs, int nx, double operand_bytes, hipStream_t st);
hipError_t ff_launch_conv(const FFPlan *pl, const FFConvSet *cs, int nc, int ngeneral, double operand_bytes,
                          hipStream_t st);
hipError_t ff_launch_inverse(const FFPlan *pl, const FFXformSet *xs, int nx, hipStream_t st);

// Work-list launches (one launch for a heterogeneous list of items).
// sh: the items' common fixed-geometry shape (ff_outer_shape), or 0.
hipError_t ff_launch_outer_list(bool inv, bool top, bool r3, int sh, const FFOuterItem *items, const uint32_t *start,
                                int n, uint32_t blocks, size_t lds_bytes, double bytes, hipStream_t st);
// The fixed-geometry kernel shape of an outer pass (top: the top pass), 0
// where it has none.
int ff_outer_shape(const FFPass *ps, bool top);
// fix12: every item's blocks are 2^12 words (the fixed-path kernel).
hipError_t ff_launch_inner_list(bool small, bool r3, bool fix12, const FFInnerItem *items, const uint32_t *start, int n,
                                uint32_t blocks, size_t lds_bytes, double bytes, hipStream_t st);
// A doubling step whose squarings are one-block transforms, fused: the CRT
// of both squares and the linear step (fib.c:36-62) in one workgroup.
// keyK / keyC must read zero before the launch.
struct FFSmallStep {
    const uint32_t *y1, *y2;  // residues of F(k)^2 and F(k-1)^2 ([3][L], CRT-ready); y2 unused if conv2 = 0
    uint32_t L1, L2;
    int bw1, bw2;
    uint32_t conv1, conv2;    // coefficients of each square (0: zero)
    uint32_t len;             // limbs of the step: n(F(k)^2) + 1, at most 3072
    int cb0;                  // 2 (-1)^k
    int bit;
    uint64_t *K, *C;          // K = bit ? B : A, C = B - A (len limbs each)
    unsigned long long *keyK, *keyC;
};
hipError_t ff_launch_small_step(const FFSmallStep *a, hipStream_t st);

// The first doubling steps of F(n), chained in one workgroup (ff_early.hip).
#define FF_EARLY_LG 10       // largest transform: 2^10 words
#define FF_EARLY_LIMBS 576   // limbs of a number in the chain
#define FF_EARLY_MAXS 16     // steps per launch
#define FF_EARLY_F0 96       // limbs of the starting pair (F(k0) for k0 < 2^13: 89 limbs)
struct FFEarlyStep {
    int lg, bw;              // transform 2^lg words, coefficient width
    uint32_t conv1;          // coefficients of the products (those of F(k)^2)
    uint32_t nw1, nw2;       // 32-bit words of F(k) and F(k-1)
    uint32_t len;            // limbs of the step's results (n(F(k)^2) + 1)
    int bit;                 // next bit of n
};
struct FFEarlyArgs {
    uint64_t f[2][FF_EARLY_F0];  // F(k0 - 1), F(k0)
    int nsteps;
    uint32_t nlo, nhi;       // limbs of the final pair to write
    FFEarlyStep st[FF_EARLY_MAXS];
    uint32_t sc[FF_EARLY_LG + 1][3];  // CRT scale of a 2^lg transform, per prime
    uint64_t *lo, *hi;       // the final pair F(k - 1), F(k)
    unsigned long long *keys;  // 2 per step: length keys of the step's new pair
};
hipError_t ff_launch_fib_early(const FFEarlyArgs *a, hipStream_t st);

// Inner blocks alone, in place (non-small plans): forward or inverse levels.
hipError_t ff_launch_inner_part_list(bool inv, const FFInnerItem *items, const uint32_t *start, int n,
                                     uint32_t blocks, size_t lds_bytes, double bytes, hipStream_t st);
size_t ff_outer_lds_bytes(const FFPass *ps);
size_t ff_inner_lds_bytes(const FFInner *in);

// ff_crt.hip
hipError_t ff_launch_crt(const FFCrtSet *js, int njobs, uint64_t max_nout, double conv_total,
                         double out_limbs_total, hipStream_t st);
// start[k]: first block of job k (blocks of 256 limbs), blocks in all.
hipError_t ff_launch_crt_list(const FFCrtJob *jobs, const uint32_t *start, int n, uint32_t blocks,
                              double conv_total, double out_limbs_total, hipStream_t st);
// Look-back words of a job of nout limbs; its workgroups in a work list.
static inline size_t ff_crt_lb_words(uint64_t nout) { return 3 + 2 * ((nout + 255) / 256); }
uint64_t ff_crt_blocks(uint64_t nout, int bw);

// A doubling step after its two squarings (k_step_fused): from the residues
// (or the limbs) of s1 = F(k)^2 and s2 = F(k-1)^2, K = bit ? F(2k+1) :
// F(2k-1) and C = F(2k) over len limbs, with their length keys.
struct FFStepJob {
    const uint32_t *y1, *y2;  // residues [3][L] of s1 and s2
    uint32_t L1, L2;
    int bw1, bw2;
    uint64_t conv1, conv2;
    const uint64_t *s1l, *s2l;  // or, when s1l is set, s1 and s2 as limbs (n1, n2 of them)
    uint64_t n1, n2;
    uint64_t len;             // limbs of K and C: n(s1) + 1
    int cb0;                  // 2 (-1)^k
    int bit;
    uint64_t *K, *C;
    unsigned long long *keyK, *keyC;
    uint32_t *lb;             // look-back workspace, ff_step_lb_words(len) zeroed words
};
size_t ff_step_lb_words(uint64_t len);
hipError_t ff_launch_step_fused(const FFStepJob *J, double conv_total, hipStream_t st);
}

#define FF_CHECK(expr)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            ff_fail(__FILE__, __LINE__, #expr, e_);                                         \
        }                                                                                   \
    } while (0)

extern "C" void ff_fail(const char *file, int line, const char *what, hipError_t e);
