// common.h -- POD structures shared by the host runtime and the gfx950 kernels.
// Kernels take one argument struct by value (hipModuleLaunchKernel kernelParams).
#pragma once
#include <stdint.h>

// Pointer members of the argument blocks point to device global memory.  The device pass
// gives them the global address space: batched kernels read their blocks from a table in
// HBM, and a pointer loaded from memory is otherwise generic, which turns every access
// through it into a flat instruction.  Host compilers see plain pointers (same layout).
#ifdef __HIP_DEVICE_COMPILE__
#define ARP_G __attribute__((address_space(1)))
#else
#define ARP_G
#endif

#ifdef __HIPCC__
#define ARP_HD __host__ __device__
#else
#define ARP_HD
struct float4_ {
    float x, y, z, w;
};
#endif

namespace arp {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr uint32_t kFlag = 0x80000000u;    // perm entry: particle has no neighbours
constexpr uint32_t kBnd = 0x80000000u;     // candidate tag: boundary particle
constexpr uint32_t kBucketCap = 128u;      // slots per reference bucket (neighbor_list.rs:45)
constexpr int kMaxSX = 2048;               // cells along x supported by the row sort
constexpr int kRowCap = 2048;              // row elements sorted in LDS (larger rows: HBM)
constexpr int kMaxK = 4;                   // stored stage derivatives for non-subdiagonal RK
constexpr int kMaxStages = 6;              // stages of a tableau (ARP_MAX_STAGES)

// mk word of a particle: material id | flags
constexpr uint32_t kMatMask = 0xFFFFu;
constexpr uint32_t kMkElastic = 1u << 16;  // solid_id valid && stiff material (fluid.rs:106)
constexpr uint32_t kMkSolid = 1u << 17;    // solid_id valid

// Material constants as the kernels read them (material.rs:11-28)
struct DMat {
    float mass, start_den, c, target_den, visc, bulk_visc;
    float lambda, mu, lambda_d, mu_d;
    float yield, hard, hard_exp, kin_hard, relax;
    uint32_t eq, plastic, elastic;
    int32_t strain_order;
    uint32_t _pad;
};
static_assert(sizeof(DMat) == 80, "DMat layout");

struct DInter {
    uint32_t potential;
    float strength, radius, dampening, friction;
};

struct GridP {
    float min[3], dim[3];
    uint32_t sub[3];
    uint32_t nrows;      // sub[1]*sub[2]
    uint32_t rowstride;  // sub[0]+1
    uint32_t dim_n;      // dimensionality (axes with dim > 0)
};

struct PhysP {
    float h, k, g, eps_hh;  // kernel radius, poly6 norm, gravity, EPSILON*h*h
    uint32_t nm;
};

// ---------------- kernel argument blocks ---------------------------------------------
struct KeyArgs {
    GridP grid;
    uint32_t n;
    const ARP_G float* posd;  // float4 per particle
    ARP_G uint32_t* keys;
    ARP_G uint32_t* ranks;
    ARP_G uint32_t* binCount;
    const ARP_G uint32_t* nlw;  // neighbour-list words (may be null): nothing to do when the stage reuses
    // when it works: copy the state (posd read above, velq) to these (may be null)
    const ARP_G float* cpyVelq;
    ARP_G float* cpyDstP;
    ARP_G float* cpyDstV;
};

constexpr int kMatLds = 16;  // material / interaction tables cached in LDS up to this many
constexpr int kScanTile = 1024;  // bins per workgroup of the row-count scan

struct ScanArgs {
    uint32_t nbins;  // rows + 1 (the last bin holds non-finite positions)
    uint32_t ntiles;
    ARP_G uint32_t* binCount;
    ARP_G uint32_t* binStart;  // nbins + 1
    ARP_G uint32_t* activeRows;
    ARP_G uint32_t* numActive;
    // decoupled look-back state: two alternating sets selected by *epoch
    ARP_G unsigned long long* tileState;  // 2 * ntiles
    ARP_G uint32_t* ticket;               // 2
    ARP_G uint32_t* epoch;                // 1
    ARP_G uint32_t* stats;                // zeroed for the following row sort
    ARP_G uint32_t* stamp;                // build counter of the target grid, advanced once per build
    ARP_G uint32_t* nlw;                  // neighbour-list words (may be null): advances their stage
                                    // counter; only the counters advance when the stage
                                    // reuses its lists
    ARP_G uint32_t* clear;                // words zeroed before the stage's kernels (may be null)
    uint32_t nclear;
};

struct ScatterArgs {
    GridP grid;
    uint32_t n;
    const ARP_G uint32_t* keys;
    const ARP_G uint32_t* ranks;
    const ARP_G uint32_t* binStart;
    ARP_G uint32_t* tmpKey;
    ARP_G uint32_t* tmpSrc;
    // non-finite particles are placed here directly (they never enter a row)
    int full;  // 1: reorder state into sorted storage; 0: only write perm
    ARP_G uint32_t* perm;
    const ARP_G float* srcPosd;
    const ARP_G float* srcVelq;
    const ARP_G uint32_t* srcMk;
    const ARP_G uint32_t* srcPid;
    ARP_G float* dstPosd;
    ARP_G float* dstVelq;
    ARP_G uint32_t* dstMk;
    ARP_G uint32_t* dstPid;
    ARP_G uint32_t* inv;  // pid -> storage index (full mode), may be null
    ARP_G uint32_t* sflag;
    ARP_G uint32_t* srcOut;  // full mode: previous storage index of every new slot (may be null)
    // neighbour-list reuse words (may be null): a stage that reuses its lists places nothing;
    // with copyOnReuse a full build then copies the state to the sorted buffers (when the
    // previous step did not leave it there)
    const ARP_G uint32_t* nlw;
    const ARP_G uint32_t* stamp;
    int copyOnReuse;
    ARP_G uint32_t* recMk;   // full build of a recording: second copy of the sorted mk / pid
    ARP_G uint32_t* recPid;
    // full build: a second state pair reordered along (velocity Verlet's force point)
    const ARP_G float* srcAuxP;
    const ARP_G float* srcAuxV;
    ARP_G float* dstAuxP;
    ARP_G float* dstAuxV;
};

struct RowSortArgs {
    GridP grid;
    uint32_t n;
    const ARP_G uint32_t* binStart;
    const ARP_G uint32_t* activeRows;
    const ARP_G uint32_t* numActive;
    const ARP_G uint32_t* tmpKey;
    const ARP_G uint32_t* tmpSrc;
    ARP_G uint32_t* gPid;  // scratch for rows larger than kRowCap
    ARP_G uint32_t* gGrp;
    ARP_G uint32_t* rowOff;       // nrows * rowstride: global start of every cell of an occupied row
    ARP_G uint32_t* rowStamp;     // nrows: build stamp of the last build that found the row occupied
    const ARP_G uint32_t* stamp;  // current build counter (advanced by the scan of this build)
    const ARP_G uint32_t* bP;     // boundary prefix per cell (ncells+1)
    const ARP_G uint32_t* rP;     // reference prefix per cell (ncells+1)
    int full;
    ARP_G uint32_t* perm;
    const ARP_G float* srcPosd;
    const ARP_G float* srcVelq;
    const ARP_G uint32_t* srcMk;
    const ARP_G uint32_t* srcPid;
    ARP_G float* dstPosd;
    ARP_G float* dstVelq;
    ARP_G uint32_t* dstMk;
    ARP_G uint32_t* dstPid;
    ARP_G uint32_t* inv;
    ARP_G uint32_t* stats;  // [0] overflowed particles (atomic), may be null
    ARP_G uint32_t* sflag;  // per storage index: cap/non-finite flag (may be null)
    ARP_G uint32_t* srcOut;
    const ARP_G uint32_t* nlw;  // neighbour-list reuse words (may be null): no row sort when reused
    ARP_G uint32_t* recMk;      // full build of a recording: second copy of the sorted mk / pid
    ARP_G uint32_t* recPid;
    const ARP_G float* srcAuxP;  // full build: a second state pair reordered along
    const ARP_G float* srcAuxV;
    ARP_G float* dstAuxP;
    ARP_G float* dstAuxV;
};

// RK stage epilogue: k -> next stage input / y_{n+1}, the b-weighted accumulator and the
// stored derivatives a non-subdiagonal tableau needs; particle registers in storage order,
// plastic-strain registers (16 floats) in solid-id order
struct RkArgs {
    const ARP_G float* Yposd;
    const ARP_G float* Yvelq;
    const ARP_G float* AccP;  // may be null
    const ARP_G float* AccV;
    const ARP_G float* Kp[kMaxK];
    const ARP_G float* Kv[kMaxK];
    float kc[kMaxK];
    int nk;
    float a_next;  // dt * a(s+1, s)
    float b_cur;   // dt * b(s)
    float b_pos;   // the same for the position part of k (velocity Verlet: 0, else b_cur)
    float vdt;     // velocity Verlet: the step (the next force point is x + dt (v + dt/2 a))
    int last;
    ARP_G float* Oposd;  // X_{s+1} or y_{n+1}
    ARP_G float* Ovelq;
    ARP_G float* Oposd2;  // second copy of the output (may be null)
    ARP_G float* Ovelq2;
    ARP_G float* AccPOut;  // may be null
    ARP_G float* AccVOut;
    ARP_G float* KpOut;  // may be null
    ARP_G float* KvOut;
    // next-stage keys
    ARP_G uint32_t* keys;
    ARP_G uint32_t* ranks;
    ARP_G uint32_t* binCount;
    // plastic strain registers
    const ARP_G float* Yeps;
    const ARP_G float* AccE;
    const ARP_G float* Ke[kMaxK];
    ARP_G float* Oeps;
    ARP_G float* AccEOut;
    ARP_G float* KeOut;
};

// Neighbour lists kept across RK stages (fluid-only scenes).  A stage whose lists are stale
// searches the grid with radius h + skin and writes, per particle, every candidate within it
// (entry k of storage index i at list[k * n + i]) and its own position; a later stage uses
// the lists while every particle stays within lim (< skin/2) of that position.  Bucket caps:
// the build counts, for every bucket, the particles within ext (>= lim) of it; no bucket can
// then hold more than that count while the lists are used, and a count above the bucket's
// free slots makes the lists unusable.  The decision is taken on the device from the stage
// counter g = words[10] (advanced by the row-count scan that opens every stage):
//   words[g & 1]     = g  when the stage at counter g ran with lists (built or reused)
//   words[2 + (g&1)] = g  when that stage found a stale entry for the next stage
//   words[4 + (g&1)] = g  when that stage built lists (its bucket counts are cleared by the
//                          next stage, from the buckets each particle recorded in cellRec)
//   words[6]         = g  of the last stage that ran with lists (read by the next stage's
//                          kernels that run before its counter advances)
//   words[8], [9]        number of list builds, number of stages run with lists
//   words[10]            the stage counter
//   words[12 + (g&1)]    stages since the lists in use at stage g were built
//   words[14]            rebuild lists older than this many stages (0: never for age)
//   words[16 + (g&1)] = g  when the stage at counter g found a particle beyond the reach of
//                          the short list (two-level lists, below)
//   words[18]            number of short-list refreshes
// Stage g reuses iff words[(g-1)&1] == g-1 and words[2+((g-1)&1)] != g-1.
// cellCnt holds two count arrays; the build at counter g counts into array g&1.
// Two-level lists (srad2 > 0): the walks use a short list of the candidates within
// h + short skin (slist/scount, the same layout), kept while every particle stays within slim
// of its position sref at the short list's build; a stage whose short list is stale walks
// the long list instead and writes a new short list from it (a refresh: no grid search).
// sref.w holds the particle's remaining reach of the long list, lim - |sref - ref|, so the one
// read of sref checks both lists.  A build writes both lists.
struct NlArgs {
    ARP_G uint32_t* words;
    ARP_G uint32_t* list;
    ARP_G uint32_t* count;
    ARP_G float* ref;         // float4 per storage index: position at the build
    ARP_G uint32_t* slist;    // two-level lists: the short list (as list), its counts and
    ARP_G uint32_t* scount;   // build positions (float4: position, remaining long-list reach)
    ARP_G float* sref;
    ARP_G uint32_t* cellCnt;  // 2 * ncells
    ARP_G uint32_t* cellRec;  // 2 per storage index: first bucket counted at the build, axis mask
    const ARP_G uint32_t* rP; // reference prefix per cell (ncells+1), with FluidArgs::bP the static entries
    uint32_t ncells;
    uint32_t cap;        // entries per particle
    uint32_t capThresh;  // counts above this are checked against the bucket's static entries
    float rad, rad2;     // h + skin and its square (build radius)
    float lim2;          // allowed squared displacement from the build position (< (skin/2)^2)
    float lim;           // its square root (negative: rebuild at every stage)
    float srad2;         // two-level lists: (h + short skin)^2 (0: one level)
    float slim2;         // allowed squared displacement from the short list's build position
    float ext;           // half-width of the box around a particle whose buckets count it
    float grey2;         // pairs farther apart than this (squared) are checked to lie in
                         // neighbouring buckets, as the reference's 27-bucket stencil requires
};

struct FluidArgs {
    GridP grid;
    PhysP phys;
    uint32_t n;
    const ARP_G DMat* mats;
    const ARP_G DInter* inter;
    const ARP_G uint32_t* perm;
    const ARP_G uint32_t* rowOff;
    const ARP_G uint32_t* rowStamp;
    const ARP_G uint32_t* stamp;  // build counter the row stamps are compared with
    int identity;           // the build reordered the state: sorted slot == storage index
    const ARP_G uint32_t* sflag;  // per storage index: no bucket slot
    const ARP_G uint32_t* stats;  // [0] particles refused by the cap in this build (may be null)
    const ARP_G uint32_t* bP;
    const ARP_G float* bposd;  // boundary (static, cell-sorted, capped)
    const ARP_G float* bvelq;
    const ARP_G uint32_t* bmk;
    // stage input (storage order)
    const ARP_G float* Xposd;
    const ARP_G float* Xvelq;
    const ARP_G uint32_t* mk;
    // 0: fused RK update; 2: write k to KdP/KdV only.  Elastic particles (when the scene
    // runs the solid path) always write their fluid part to KdP/KdV.
    int mode;
    int skip_elastic;
    ARP_G float* KdP;
    ARP_G float* KdV;
    RkArgs rk;
    // optional record of every particle's neighbours within h (for the reverse sweep):
    // entry k of storage index i at nbList[k * n + i], count at nbCount[i]; a particle with
    // more than nbCap neighbours sets *nbOver
    ARP_G uint32_t* nbList;
    ARP_G uint32_t* nbCount;
    ARP_G uint32_t* nbOver;
    uint32_t nbCap;
    NlArgs nl;  // neighbour-list reuse (k_fluid_nl, k_fluid_nlb and their recording forms)
    ARP_G uint32_t* recUse;  // recording with lists: set to whether the stage walked kept lists
    // a list build after a full reorder also writes the sorted mk / pid into the buffers the
    // reorder read (so that both copies agree and a reusing step needs no reorder)
    ARP_G uint32_t* mkMirror;
    ARP_G uint32_t* pidMirror;
    const ARP_G uint32_t* pidIn;
    // one mobile material and one boundary material, no elastic particles: the list walk
    // takes the materials of a pair from these (k_fluid_nl_f, k_fluid_nlr_f)
    uint32_t fastMat;
    uint32_t fastBnd;
    // speculative walk (k_fluid_nl(_f), k_fluid_nlr(_f) of a list-reusing step issued without
    // the rebuild chain): the stage keeps the chain's bookkeeping itself (the stage counter,
    // the build stamp, the refused count) and returns at once, recording specStep in the list
    // words and in *hostAbort, when the lists are not valid for it; later stages then return
    // at once
    uint32_t spec;
    uint32_t specGen;     // the stage's list counter value
    uint32_t specStep;    // (step << 4) | stage
    ARP_G uint32_t* specSnap;   // first stage of a step: words[0..14], *stampW, statsW[0] saved here
    ARP_G uint32_t* hostAbort;  // host-mapped
    ARP_G uint32_t* stampW;
    ARP_G uint32_t* statsW;
    // first stage of a speculative recorded step: the step's record flags to clear
    // (kRecFlags words) and its storage-order records (identity source index, mk,
    // pid), which the rebuild chain's reusing sort writes otherwise
    ARP_G uint32_t* specClear;
    ARP_G uint32_t* recSrc;
    ARP_G uint32_t* recMk;
    ARP_G uint32_t* recPid;
};

// words of the neighbour-list state beyond the decision words
constexpr int kNlAbort = 15;  // the aborting speculative step (0xFFFFFFFF: none)
constexpr int kNlWords = 32;
constexpr int kSpecSnap = 24;  // words per speculative step snapshot
constexpr int kRecFlags = 12;  // a recorded step's flags: overflow and list use per stage (2 * 6 stages)

// Restore of the list words after an aborted speculative step (one thread).
struct SpecRestoreArgs {
    ARP_G uint32_t* words;
    const ARP_G uint32_t* snap;  // words[0..14], stamp, stats (null: only the abort word and the counter)
    ARP_G uint32_t* stamp;
    ARP_G uint32_t* stats;
    // a recorded step keeps one build stamp per stage: those of its later stages that ran
    // (each advanced by one) go back by one
    ARP_G uint32_t* stampDec;
    uint32_t ndec;
};

// Solid path (strain.rs, solid.rs): static per-solid reference-neighbour lists
struct SolidArgs {
    GridP grid;
    PhysP phys;
    uint32_t ns;
    const ARP_G DMat* mats;
    const ARP_G uint32_t* part;     // particle id per solid
    const ARP_G uint32_t* inv;      // particle id -> storage index
    const ARP_G float* ref;         // float4 reference position per solid
    const ARP_G float* Linv;        // corrected-gradient L^-1 per solid, 16 (column-major)
    const ARP_G uint32_t* sOff;     // strain list: ns+1 offsets
    const ARP_G uint32_t* sPid;     //   neighbour particle id
    const ARP_G float* sG;          //   float4 (L^-1 gradW(dX) xyz, V2)
    const ARP_G uint32_t* sSid;     //   internal index of the entry's solid (static)
    const ARP_G float* sG2;         //   float4 (L_s2^-1 gradW(X_s - X_s2) xyz, 0): the weight of this
                                    //   solid in neighbour s2's F (the reverse pass; static)
    const ARP_G uint32_t* eOff;     // elastic-force list (same material): ns+1 offsets
    const ARP_G uint32_t* ePid;     //   neighbour particle id
    const ARP_G uint32_t* eSid;     //   neighbour solid id
    const ARP_G float* eG;          //   float4 (gradW(dX) xyz, W(dX))
    const ARP_G float* Xposd;       // stage input, storage order
    const ARP_G float* Xvelq;
    const ARP_G uint32_t* mk;
    const ARP_G float* Xeps;        // stage input plastic strain per solid (16)
    // when set: y_n's plastic strain, which a solid of a non-plastic material keeps through
    // the step (its rate is zero): its strain reads it here and its RK update of the plastic
    // strain is skipped (null: every solid updates it)
    const ARP_G float* epsY;
    ARP_G float* A;                 // per solid: (P L^-1), three rows of three floats
    ARP_G float* Fm;                // per solid: F, three rows of three floats
    ARP_G float* Fd;                // per solid: F', three rows of three floats (recorded stages; may be null)
    ARP_G float* keps;              // per solid: plastic strain rate (16)
    // current grid, for the hourglass stencil test and self contact
    const ARP_G uint32_t* perm;
    const ARP_G uint32_t* rowOff;
    const ARP_G uint32_t* rowStamp;
    const ARP_G uint32_t* stamp;  // build counter the row stamps are compared with
    const ARP_G uint32_t* sflag;    // per storage index: 1 when refused by the cap / non-finite
    const ARP_G uint32_t* pid;      // storage -> particle id
    const ARP_G uint32_t* sidByPid;
    // per grid row, written by k_strain and read by the self-contact sweeps: min and max of
    // u = x - X over the row's solids (.w of the first: 1 when one lies outside the grid)
    ARP_G float* rowBox;
    const ARP_G float* kscale;      // per solid stiffness scale of (lambda, mu), or null (1)
    // fluid part of elastic particles (written by k_fluid)
    const ARP_G float* KFp;
    const ARP_G float* KFv;
    int mode;                 // 0: RK update; 2: write k to KdP/KdV and keps only
    ARP_G float* KdP;
    ARP_G float* KdV;
    RkArgs rk;
};

// Adjoint of the solid part of one stage: pass 1 per solid gathers the cotangents of its
// stress (elastic force, plastic rate), its F (hourglass) and the direct position / density
// terms, then runs the strain adjoint (strain orders other than 2 complete it in a pass of
// their own); pass 2 turns the F, F' cotangents of a solid and its reference neighbours into
// position / velocity cotangents.
struct SolidAdjArgs {
    SolidArgs f;        // the stage's forward arguments (A, Fm recomputed by k_strain)
    const ARP_G float* Wp;    // stage cotangent of k, times wscale
    const ARP_G float* Wv;
    float wscale;
    ARP_G float* Gp;          // += solid part of dL/dX
    ARP_G float* Gv;
    ARP_G float* LamP;
    ARP_G float* LamV;
    ARP_G float* Fbar;        // per solid 18: cotangent rows of F (3x3) then of F' (3x3)
    ARP_G float* spar;        // per solid 4 (+=): dL/d (lambda, mu, lambda_d, mu_d) as the solid sees them
    // per solid of a material with strain_order != 2: the cotangents of E and E' (16 each,
    // column-major) and F' (9, row-major), 48 floats, completed into Fbar by
    // k_solid_adj_strain (null when the scene has no such material)
    ARP_G float* Ebuf;
    // plastic flow (null lamE: no plastic material).  Per solid 16 floats each (column-major):
    // lamE = dL/d (the step's plastic strain), gE[t] = dL/d (stage t's input plastic strain).
    // The stage's plastic-rate cotangent is wcl lamE + sum over t > stage of wcg[t] gE[t];
    // the last reverse stage (stage 0) adds every gE[t] to lamE (dL/d y_n).
    ARP_G float* lamE;
    ARP_G float* gE[kMaxStages];
    float wcl;
    float wcg[kMaxStages];
    int stage, nstages;
};

struct InvArgs {
    uint32_t n;
    const ARP_G uint32_t* pid;
    ARP_G uint32_t* inv;
};

// Adjoint of one fluid stage: dL/dX_s = J_f(X_s)^T w_s (gather over the stage's neighbour
// structure), accumulated into the step cotangent, plus per-material parameter partials.
struct AdjArgs {
    GridP grid;
    PhysP phys;
    uint32_t n;
    const ARP_G DMat* mats;
    const ARP_G DInter* inter;
    const ARP_G uint32_t* perm;
    const ARP_G uint32_t* rowOff;
    const ARP_G uint32_t* rowStamp;
    const ARP_G uint32_t* stamp;  // build counter the row stamps are compared with
    int identity;
    const ARP_G uint32_t* sflag;
    const ARP_G uint32_t* bP;
    const ARP_G float* bposd;
    const ARP_G float* bvelq;
    const ARP_G uint32_t* bmk;
    const ARP_G float* Xposd;
    const ARP_G float* Xvelq;
    const ARP_G uint32_t* mk;
    const ARP_G float* Wp;   // cotangent of k: wscale * (pos'.xyz, den')
    const ARP_G float* Wv;   // wscale * (vel'.xyz, 0)
    float wscale;
    ARP_G float* Gp;         // dL/dX: (x.xyz, rho)
    ARP_G float* Gv;         // (v.xyz, 0)
    ARP_G float* LamP;       // step cotangent accumulator (+= G)
    ARP_G float* LamV;
    ARP_G float* pgrad;      // per block: nm * 2 (d bulk_visc, d visc), accumulated over the sweep
    // neighbours recorded by the forward stage (null, or *nbOver set: search again)
    const ARP_G uint32_t* nbList;
    const ARP_G uint32_t* nbCount;
    const ARP_G uint32_t* nbOver;
    uint32_t fastMat;  // the mobile particles' material (kernels without the _g suffix)
    // recorded with kept lists: nonzero when the forward stage walked them (it then kept
    // the storage order and sorted nothing: threads map to storage indices one to one)
    const ARP_G uint32_t* stageUse;
    // first reverse stage of a step: the accumulator starts from these (else LamP/LamV)
    const ARP_G float* LamInP;
    const ARP_G float* LamInV;
    // with LamInP: also form the next reverse stage's W = wl * Lam_in + wg * G here
    ARP_G float* WnP;
    ARP_G float* WnV;
    float wl, wg;
    // last reverse stage of a step (no solid passes after it): the accumulator goes to
    // LamOut[permOut[i]] (the storage order before the step) instead of LamP
    const ARP_G uint32_t* permOut;
    ARP_G float* LamOutP;
    ARP_G float* LamOutV;
};

// W = cl * lambda + sum_t c_t * G_t (stage cotangent of k)
struct AdjWArgs {
    uint32_t n;
    float cl;
    const ARP_G float* LP;
    const ARP_G float* LV;
    int ng;
    const ARP_G float* GP[kMaxK + 2];
    const ARP_G float* GV[kMaxK + 2];
    float cg[kMaxK + 2];
    ARP_G float* WP;
    ARP_G float* WV;
};

// out[src[t]] = in[t] (cotangent back to the previous storage order), or out = in
struct AdjPermArgs {
    uint32_t n;
    const ARP_G uint32_t* src;
    const ARP_G float* inP;
    const ARP_G float* inV;
    ARP_G float* outP;
    ARP_G float* outV;
};

// cotangent AoS (ArpParticle records holding dL/d(den, pos, vel)) <-> SoA in storage order
struct GradAosArgs {
    uint32_t n;
    int to_aos;
    const ARP_G uint32_t* pid;
    ARP_G void* aos;
    ARP_G float* P;
    ARP_G float* V;
};

struct ReduceArgs {
    uint32_t nblocks, width;
    const ARP_G float* part;  // nblocks * width
    ARP_G float* out;         // width, accumulated (+=)
};

// A batch of independent scenes in one launch: block L belongs to scene blockScene[L] and
// is block L - blockFirst[s] of blockFirst[s+1] - blockFirst[s] blocks of that scene; the
// scene's argument block is args + s * stride.
struct BatchArgs {
    const ARP_G void* args;
    uint32_t stride;
    const ARP_G uint32_t* blockScene;
    const ARP_G uint32_t* blockFirst;  // nscenes + 1
};

struct AosInArgs {
    uint32_t n;
    const ARP_G void* aos;  // ArpParticle*
    const ARP_G DMat* mats;
    const ARP_G uint32_t* mkStatic;  // per pid
    ARP_G float* posd;
    ARP_G float* velq;
    ARP_G uint32_t* mk;
    ARP_G uint32_t* pid;
};

struct AosOutArgs {
    uint32_t n;
    const ARP_G float* posd;
    const ARP_G float* velq;
    const ARP_G uint32_t* pid;
    const ARP_G uint32_t* sidByPid;
    const ARP_G uint32_t* matByPid;
    ARP_G void* aos;
};

struct LincombArgs {
    uint32_t n, ns;
    uint32_t nterms;
    float coef[8];
    const ARP_G void* terms[8];   // ArpParticle*
    const ARP_G void* sterms[8];  // ArpSolidParticle*
    ARP_G void* out;
    ARP_G void* sout;
};

// Velocity Verlet before the force evaluation of a step, from y and the last evaluation
// D = (a, rho', eps'): in place Z = (x + dt (v + dt/2 a), v + dt/2 a, rho + dt/2 rho',
// eps + dt/2 eps') (the base of the step's kick), and the force point
// X = (Z.x, Z.v + dt/2 a, Z.rho + dt/2 rho', Z.eps + dt/2 eps'); q = p/rho^2 of each.
struct VerletArgs {
    uint32_t n, ns;
    float dt;
    ARP_G float* posd;        // y -> Z
    ARP_G float* velq;
    ARP_G float* Xposd;       // X
    ARP_G float* Xvelq;
    const ARP_G float* Dp;    // the last force evaluation: (v, rho')
    const ARP_G float* Dv;    // (a, 0)
    const ARP_G uint32_t* mk;
    const ARP_G DMat* mats;
    ARP_G float* eps;         // plastic strain registers (16 per solid), may be null: -> Z
    ARP_G float* Xeps;        // -> X
    const ARP_G float* De;    // their last rates
};

struct FillArgs {
    ARP_G uint32_t* p;
    uint32_t n, v;
};

struct EpsInArgs {  // AoS solids -> plastic strain registers
    uint32_t ns;
    const ARP_G void* aos;
    ARP_G float* eps;
    const ARP_G uint32_t* order;  // internal solid -> its record in aos
};

struct EpsOutArgs {  // plastic strain (or its rate) -> AoS solids {part_id, ref_pos, stress}
    uint32_t ns;
    const ARP_G float* eps;
    const ARP_G float* ref;  // null: write ref_pos = 0 (derivative)
    const ARP_G uint32_t* part;
    ARP_G void* aos;
    const ARP_G uint32_t* order;  // internal solid -> its record in aos
};

struct ClearSolidsArgs {
    uint32_t ns;
    const ARP_G void* in;  // ArpSolidParticle*
    ARP_G void* out;
};

struct VelocityArgs {
    uint32_t n, ns;
    float r;
    const ARP_G void* in;
    const ARP_G void* sin;
    ARP_G void* out;
    ARP_G void* sout;
};

}  // namespace arp
