// runtime.cpp -- host runtime behind include/arp.h: context, scene upload (FluidSim::new),
// static boundary/reference grids, the per-stage launch chain and the fused RK step.
// Kernels come from the embedded gfx950 code object (hipModuleLoadData +
// hipModuleLaunchKernel); no CUDA-compat layer.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/arp.h"
#include "common.h"
#include "scratch.h"

using namespace arp;

extern "C" const unsigned char arp_kernels_co[];

// HIP's extended module launch (hip/hip_ext.h; that header needs the device-side headers):
// the start/stop events are recorded by the dispatch itself.
extern "C" hipError_t hipExtModuleLaunchKernel(hipFunction_t f, uint32_t globalWorkSizeX, uint32_t globalWorkSizeY,
                                               uint32_t globalWorkSizeZ, uint32_t localWorkSizeX,
                                               uint32_t localWorkSizeY, uint32_t localWorkSizeZ,
                                               size_t sharedMemBytes, hipStream_t hStream, void** kernelParams,
                                               void** extra, hipEvent_t startEvent, hipEvent_t stopEvent,
                                               uint32_t flags);

namespace {

enum Kern {
    K_KEY,
    K_SCAN,
    K_SCATTER,
    K_ROWSORT,
    K_FLUID,
    K_AOS_IN,
    K_AOS_OUT,
    K_LINCOMB,
    K_VELOCITY,
    K_FILL,
    K_CLEAR_SOLIDS,
    K_STRAIN,
    K_SOLID,
    K_EPS_IN,
    K_EPS_OUT,
    K_FLUID_ADJ,
    K_ADJ_W,
    K_ADJ_PERM,
    K_GRAD_AOS,
    K_REDUCE,
    K_KEY_B,
    K_SCAN_B,
    K_SCATTER_B,
    K_ROWSORT_B,
    K_FLUID_B,
    K_INV,
    K_SOLID_ADJ1,
    K_SOLID_ADJ2,
    K_FLUID_REC,
    K_FLUID_ADJ_REC,
    K_FLUID_NL,
    K_FLUID_NL_B,
    K_FLUID_NLB,
    K_FLUID_NLB_B,
    K_FLUID_ADJ_G,
    K_FLUID_ADJ_REC_G,
    K_FLUID_NLBR,
    K_FLUID_NLR,
    K_FLUID_ADJ_F,
    K_FLUID_ADJ_REC_F,
    K_VERLET_PRE,
    K_FLUID_NL_F,
    K_FLUID_NLR_F,
    K_SPEC_RESTORE,
    K_FLUID_NL_F_B,
    K_STRAIN_B,
    K_SOLID_B,
    K_FLUID_NL_F_B_BUF,
    K_FLUID_NL_B_BUF,
    K_SOLID_ADJ_STRAIN,
    K_FLUID_NLBW_F_B,
    K_COUNT
};
const char* kKernNames[K_COUNT] = {"k_key",     "k_scan",    "k_scatter",  "k_rowsort",
                                   "k_fluid",   "k_aos_in",  "k_aos_out",  "k_lincomb",
                                   "k_velocity", "k_fill",   "k_clear_solids", "k_strain",
                                   "k_solid",   "k_eps_in",  "k_eps_out", "k_fluid_adj",
                                   "k_adj_w",   "k_adj_perm", "k_grad_aos", "k_reduce",
                                   "k_key_batch", "k_scan_batch", "k_scatter_batch", "k_rowsort_batch",
                                   "k_fluid_batch", "k_inv", "k_solid_adj1", "k_solid_adj2",
                                   "k_fluid_rec", "k_fluid_adj_rec", "k_fluid_nl", "k_fluid_nl_batch",
                                   "k_fluid_nlb", "k_fluid_nlb_batch",
                                   "k_fluid_adj_g", "k_fluid_adj_rec_g", "k_fluid_nlbr", "k_fluid_nlr",
                                   "k_fluid_adj_f", "k_fluid_adj_rec_f", "k_verlet_pre", "k_fluid_nl_f", "k_fluid_nlr_f",
                                   "k_spec_restore", "k_fluid_nl_f_batch", "k_strain_batch", "k_solid_batch",
                                   "k_fluid_nl_f_batch_buf", "k_fluid_nl_batch_buf", "k_solid_adj_strain",
                                   "k_fluid_nlbw_f_batch"};

constexpr size_t kPad = 16;  // elements of slack after every per-particle array

// The list walks form 32-bit byte offsets of entries (entry k of particle i lies at element
// k * (n + kPad) + i), so every list entry must lie below 4 GiB.
static bool nl_fits_u32(uint32_t cap, uint32_t n) {
    return (uint64_t)cap * ((uint64_t)n + kPad) * 4u < (1ull << 32);
}
constexpr uint32_t kRecNb = 32;  // neighbours per particle and stage kept by a recording

struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

template <class T>
T* P(DBuf& b) {
    return reinterpret_cast<T*>(b.p);
}

// Where a neighbour build writes its structure (the context's own buffers, or a recording
// slot kept for the adjoint of that stage).
struct GridSlot {
    uint32_t* perm = nullptr;
    uint32_t* sflag = nullptr;
    uint32_t* rowOff = nullptr;
    uint32_t* rowStamp = nullptr;
    uint32_t* stamp = nullptr;  // device build counter of this slot (rows stamped with it are current)
    bool full = false;
};

// One recorded step of a checkpoint segment: everything the stage adjoints read.
static_assert(kRecFlags == 2 * ARP_MAX_STAGES, "recorded step flags: two per stage");
static_assert(kMaxStages == ARP_MAX_STAGES, "stage limit of the kernels' argument blocks");

struct RecStep {
    DBuf mk, pid, src, stamps;
    uint32_t* fl = nullptr;  // nbOver[ARP_MAX_STAGES], use[ARP_MAX_STAGES] (in the context's recFlags)
    uint32_t over[ARP_MAX_STAGES] = {};  // host copy of nbOver after the recording
    DBuf nbList[ARP_MAX_STAGES], nbCount[ARP_MAX_STAGES];
    DBuf Xp[ARP_MAX_STAGES], Xv[ARP_MAX_STAGES];
    DBuf perm[ARP_MAX_STAGES], sflag[ARP_MAX_STAGES], rowOff[ARP_MAX_STAGES], rowStamp[ARP_MAX_STAGES];
    // solid stages: the strain kernel's A = P L^-1, F and self-contact row boxes, written by
    // the recorded stage itself (so that the reverse stage does not recompute them)
    DBuf sA[ARP_MAX_STAGES], sF[ARP_MAX_STAGES], sFd[ARP_MAX_STAGES], sBox[ARP_MAX_STAGES];
    DBuf sEps[ARP_MAX_STAGES];  // each stage's input plastic strain (scenes with plastic flow)
    GridSlot slot[ARP_MAX_STAGES];
};

struct Checkpoint {
    DBuf P, V, mk, pid;
    DBuf E;  // the plastic strain (scenes with a plastic material)
};

// A launch recorded instead of issued (the batch path collects one per scene and issues
// them together).
struct Launch {
    int k;
    uint32_t blocks, threads;
    std::vector<uint8_t> args;
};

}  // namespace

struct arp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    hipModule_t mod = nullptr;
    hipFunction_t fn[K_COUNT] = {};
    std::string err;
    std::vector<DBuf*> bufs;

    // scene
    bool have_scene = false;
    uint32_t n = 0, ns = 0, nb = 0, nm = 0;
    uint64_t ncells = 0;
    GridP grid{};
    PhysP phys{};
    std::vector<ArpMaterial> mats_h;
    std::vector<uint32_t> mat_by_pid, sid_by_pid, mk_by_pid;
    // Solids are numbered internally in the cell order of their reference positions, so that
    // a thread per solid reads its particle's state (stored cell-sorted) coalesced:
    // s_ord[t] = external solid id of internal solid t, s_rank = the inverse.  Every per-solid
    // device array is in the internal order; the solid records the caller sees are not.
    std::vector<uint32_t> s_ord, s_rank;
    int64_t mobile_mat = -1;  // the material of every mobile particle, or -1 (none or several)
    int64_t bnd_mat = -1;     // the material of every boundary particle (none: -1, several: -2)
    bool bnd_elastic = false; // some boundary word carries the elastic flag
    bool any_elastic = false; // some particle is elastic (a solid of a stiff material)
    std::vector<ArpSolidParticle> solids_h;
    std::vector<ArpParticle> boundary_h;    // the uploaded boundary, interactions and bounds
    std::vector<ArpMatInteraction> inter_h; // (kept for arp_add_particles)
    ArpAABB bounds_h{};
    float alpha_h = 0.0f;
    std::vector<float> kscale_h;  // per-solid stiffness scale (empty: 1)
    std::vector<uint32_t> b_orig;  // sorted boundary slot -> original boundary id
    std::vector<uint32_t> bP_h, rP_h;

    DBuf mats, inter, bposd, bvelq, bmk, bP, rP, mkStatic, matByPid, sidByPid, sidExtByPid, sOrd;
    // state
    DBuf Ysort[2], Yland[2];  // [0] posd, [1] velq
    // the live state is only in Ysort (same storage order): a list-reusing RK step of several
    // stages leaves y_{n+1} there alone; sync_yland copies it back before Yland is read
    bool yland_stale = false;
    DBuf meta_mk[2], meta_pid[2];
    int cur = 0;
    DBuf Xb[2][2];  // [buffer][posd/velq]
    DBuf Acc[2];
    DBuf K[kMaxK][2];
    DBuf Dv[2], De;            // velocity Verlet: the last force evaluation (storage order)
    DBuf XF[2][2], XFe;        // velocity Verlet: the force point, [landing / sorted][posd / velq]
    bool verlet_valid = false; // Dv/De hold the derivative at the live state's last force point
    DBuf inv, sflag, KF[2];
    // solids: static reference-neighbour data and plastic-strain registers (solid-id order)
    bool solids_active = false;
    DBuf part_d, ref_d, Linv_d, sOff_d, sPid_d, sSid_d, sG_d, sG2_d, eOff_d, ePid_d, eSid_d, eG_d;
    DBuf Yeps, Xeps[2], AccE, Ke[kMaxK], keps, A_d, Fm_d, ps_eps, rowBox;
    // per-stage API buffers
    DBuf ps_mk[2], ps_pid[2], aos_tmp, ps_keys, saos_tmp;
    // grid
    DBuf keys, ranks, binCount, binStart, activeRows, numActive, tmpKey, tmpSrc, perm, rowOff,
        gPid, gGrp, stats, scanState, scanTicket, scanEpoch, rowStamp;
    DBuf stampDev;          // build counter written into rowStamp by the row sort
    GridSlot slot;          // where builds go (the buffers above unless recording)
    // neighbour lists kept across stages (fluid-only scenes, the plain step)
    float nl_skin = 0.2f;   // list radius h * (1 + skin); 0 disables
    float nl_sskin = 0.0f;  // two-level lists: the short list's skin (0: one level)
    uint32_t nl_cap = 48;   // entries per particle
    bool nl_always = false; // rebuild at every stage (testing)
    bool nl_nospec = false; // issue the rebuild chain at every stage (no speculative steps)
    // speculative steps: host-mapped abort word, per-step snapshots of the list words, and
    // the events of the steps in flight
    uint32_t* hAbort = nullptr;
    uint32_t* hAbortDev = nullptr;
    DBuf nlSnap;
    hipEvent_t specEv[4] = {};
    uint64_t spec_aborts = 0;
    uint64_t spec_steps = 0;  // speculative steps that completed
    uint32_t nl_max_age = 0;  // rebuild lists older than this many stages (0: never for age)
    bool nl_dirty = true;   // the decision words and bucket counts must be reset
    int64_t nl_thresh = -1; // bucket counts above this are checked exactly (-1: not derived)
    bool nl_suspend = false;  // a batch whose scenes cannot all keep lists runs without
    // recording: the step recorded after the current one (its stage input can be written by
    // the current step), and the stage-input buffer that already holds the live state
    RecStep* rec_next = nullptr;
    const void* rec_prefilled = nullptr;
    DBuf nlWords, nlList, nlCount, nlRef, nlCell, nlCellRec, nlSList, nlSCount, nlSRef;
    // adjoint recording
    // this GPU's fixed scratch array (scratch.hip) for the entry points that allocate nothing:
    // its device address, its size, and the HIP error of the lookup (0: found)
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    int scratch_err = 0;
    bool rec_valid = false;
    uint32_t rec_nb_cap = kRecNb;  // neighbours per particle and stage of the recording
    float rec_dt = 0.0f;
    uint32_t rec_nsteps = 0, rec_K = 0, rec_seg = 0;
    ArpTableau rec_tab{};
    std::vector<Checkpoint> ck;
    std::vector<RecStep> rec;
    DBuf LamIn[2], LamAcc[2], LamNext[2], Wb[2], Gb[ARP_MAX_STAGES][2], pgradPart, pgradSum, gaos;
    DBuf Fbar, spar;  // solid reverse mode: per-solid F cotangents, per-solid stiffness partials
    DBuf Ebuf;        // solid reverse mode, strain orders != 2: E, E' cotangents and F' per solid
    DBuf LamE, GbE[ARP_MAX_STAGES];  // plastic flow: the plastic strain's step and stage cotangents
    DBuf kscale;      // per-solid stiffness scale (empty: 1)
    DBuf recFlags;    // the recorded steps' stage flags, 2 * ARP_MAX_STAGES words per step
    bool keys_valid = false;
    // what the last build described (for the dump)
    const DBuf* last_pid = nullptr;
    const DBuf* last_posd = nullptr;
    bool have_build = false;

    // profiling: event pairs around launches
    bool prof = false;
    bool prof_sel[K_COUNT];  // kernels timed while profiling
    std::vector<Launch>* capture = nullptr;
    bool capture_solid = false;  // a batch holds scenes with solids: emit the solid path's launches  // record launches here instead of issuing them
    bool capture_fm = false;  // a capture may take the fast-material walk (every scene of its batch can)
    uint32_t rowsort_div = 32;               // particles per row-sort workgroup (tuning knob)
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, size_t>> ev_marks;  // (kernel, index of first event)

    // integrator
    ArpTableau tab{};
    float timestep = 0.01f;
    uint32_t subticks = 1;
    double time = 0.0;

    arp_ctx() {
        for (bool& b : prof_sel) b = true;
        bufs = {&mats,     &inter,     &bposd,      &bvelq,    &bmk,    &bP,       &rP,
                &mkStatic, &matByPid,  &sidByPid,   &sidExtByPid, &sOrd, &Ysort[0], &Ysort[1], &Yland[0], &Yland[1],
                &meta_mk[0], &meta_mk[1], &meta_pid[0], &meta_pid[1], &Xb[0][0], &Xb[0][1],
                &Xb[1][0], &Xb[1][1],  &Acc[0],     &Acc[1],   &inv,    &ps_mk[0], &ps_mk[1],
                &ps_pid[0], &ps_pid[1], &aos_tmp,   &ps_keys,  &keys,   &ranks,    &binCount,
                &binStart, &activeRows, &numActive, &tmpKey,   &tmpSrc, &perm,     &rowOff,
                &gPid,     &gGrp,      &stats,    &scanState, &scanTicket, &scanEpoch, &rowStamp,
                &stampDev, &nlWords,   &nlList,   &nlCount,  &nlRef,  &nlCell, &nlCellRec, &nlSList, &nlSCount, &nlSRef,
                &Dv[0],    &Dv[1],     &De,       &XF[0][0], &XF[0][1], &XF[1][0], &XF[1][1], &XFe};
        for (int k = 0; k < kMaxK; k++) {
            bufs.push_back(&K[k][0]);
            bufs.push_back(&K[k][1]);
            bufs.push_back(&Ke[k]);
        }
        for (DBuf* b : {&sflag, &KF[0], &KF[1], &part_d, &ref_d, &Linv_d, &sOff_d, &sPid_d, &sSid_d, &sG_d, &sG2_d, &eOff_d,
                        &ePid_d, &eSid_d, &eG_d, &Yeps, &Xeps[0], &Xeps[1], &AccE, &keps, &A_d, &Fm_d, &ps_eps, &rowBox,
                        &saos_tmp})
            bufs.push_back(b);
    }
};

namespace {

int fail(arp_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, x)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail(c, ARP_E_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_),  \
                        __FILE__, __LINE__);                                               \
    } while (0)

int dalloc(arp_ctx* c, DBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.p && b.bytes >= bytes) return ARP_OK;
    if (b.p) {
        hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) return fail(c, ARP_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    b.bytes = bytes;
    return ARP_OK;
}

int upload(arp_ctx* c, DBuf& b, const void* h, size_t bytes) {
    int rc = dalloc(c, b, bytes);
    if (rc) return rc;
    if (bytes) HIPCHK(c, hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, c->stream));
    return ARP_OK;
}

hipEvent_t next_event(arp_ctx* c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

template <class A>
int launch(arp_ctx* c, Kern k, uint32_t blocks, uint32_t threads, const A& args) {
    if (c->capture) {  // (a launch of no workgroups too: the scene takes no part in it)
        Launch L{(int)k, blocks, threads, std::vector<uint8_t>(sizeof(A))};
        memcpy(L.args.data(), &args, sizeof(A));
        c->capture->push_back(std::move(L));
        return ARP_OK;
    }
    if (blocks == 0) return ARP_OK;
    void* params[] = {const_cast<A*>(&args)};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->prof && c->prof_sel[k]) {
        size_t first = c->ev_used;
        e0 = next_event(c);
        e1 = next_event(c);
        if (e0 && e1) c->ev_marks.push_back({(int)k, first});
    }
    hipError_t e;
    if (e0 && e1)  // the events time the dispatch itself (no marker packets around it)
        e = hipExtModuleLaunchKernel(c->fn[k], blocks * threads, 1, 1, threads, 1, 1, 0, c->stream, params,
                                     nullptr, e0, e1, 0);
    else
        e = hipModuleLaunchKernel(c->fn[k], blocks, 1, 1, threads, 1, 1, 0, c->stream, params, nullptr);
    if (e != hipSuccess)
        return fail(c, ARP_E_HIP, "launch %s failed: %s", kKernNames[k], hipGetErrorString(e));
    return ARP_OK;
}

inline uint32_t nblk(uint64_t n, uint32_t b = 256) { return (uint32_t)((n + b - 1) / b); }
// workgroups of a grid-stride pass that a stage reusing its lists skips (its launch costs
// about a workgroup dispatch each): four blocks of particles per workgroup, at most 1024
inline uint32_t gated_blocks(uint64_t n) { return std::min<uint32_t>(std::max<uint32_t>(nblk(n) / 4, 1), 1024); }

int fill_u32(arp_ctx* c, DBuf& b, uint32_t n, uint32_t v) {
    FillArgs a{P<uint32_t>(b), n, v};
    return launch(c, K_FILL, nblk(n), 256, a);
}

// ---- host restatements of FluidSim::new derivations --------------------------------
float vol_nball(uint32_t n, float r) {
    if (n == 0) return 1.0f;
    if (n == 1) return 2.0f * r;
    const float pi = 3.14159265358979323846f;
    return 2.0f * pi * r * r * vol_nball(n - 2, r) / (float)n;
}

// host copy of the kernel's cell function (identical IEEE float operations)
bool host_cell(const GridP& g, const float* p, uint32_t& cx, uint32_t& cy, uint32_t& cz) {
    for (int i = 0; i < 3; i++)
        if (!std::isfinite(p[i])) return false;
    uint32_t c[3];
    for (int i = 0; i < 3; i++) {
        float t = (p[i] - g.min[i]) / g.dim[i];
        t = std::fmin(std::fmax(t, 0.0f), 1.0f);
        float f = std::floor(t * (float)g.sub[i]);
        uint32_t v = (uint32_t)f;
        c[i] = v >= g.sub[i] ? g.sub[i] - 1 : v;
    }
    cx = c[0];
    cy = c[1];
    cz = c[2];
    return true;
}

float host_pressure(const ArpMaterial& m, float den) {
    float c = m.sound_speed, d0 = m.target_den, p;
    switch (m.state_eq) {
        case 1: p = c * c; break;
        case 2: p = c * c * (den - d0); break;
        case 3: p = (c * c * d0 / 7.0f) * (std::pow(den / d0, 7.0f) - 1.0f); break;
        default: p = 0.0f;
    }
    return p / (den * den);
}

bool is_stiff(const ArpMaterial& m) { return m.normal_stiffness != 0.0f || m.shear_stiffness != 0.0f; }

// an elastic material whose Seth-Hill order is not 2 (its strain adjoint needs
// k_solid_adj_strain)
static bool any_other_order(const arp_ctx* c) {
    for (uint32_t m = 0; m < c->nm; m++)
        if (is_stiff(c->mats_h[m]) && c->mats_h[m].strain_order != 2) return true;
    return false;
}

// an elastic material with plastic flow (its plastic strain is state the reverse sweep
// carries)
static bool any_plastic(const arp_ctx* c) {
    for (uint32_t m = 0; m < c->nm; m++)
        if (is_stiff(c->mats_h[m]) && c->mats_h[m].plastic) return true;
    return false;
}

// stable counting sort of items by cell with a per-cell cap; returns the prefix over cells
// of kept items and the kept items in (cell, id) order
void cell_sort(uint64_t ncells, const std::vector<uint64_t>& key, const std::vector<uint32_t>& cap_room,
               std::vector<uint32_t>& prefix, std::vector<uint32_t>& kept) {
    std::vector<uint32_t> cnt(ncells + 1, 0);
    for (uint64_t k : key)
        if (k != UINT64_MAX) cnt[k + 1]++;
    std::vector<uint32_t> start(ncells + 1, 0);
    for (uint64_t c = 0; c < ncells; c++) start[c + 1] = start[c] + cnt[c + 1];
    std::vector<uint32_t> order(start[ncells]);
    std::vector<uint32_t> fillc(ncells, 0);
    for (size_t i = 0; i < key.size(); i++)
        if (key[i] != UINT64_MAX) order[start[key[i]] + fillc[key[i]]++] = (uint32_t)i;
    prefix.assign(ncells + 1, 0);
    kept.clear();
    for (uint64_t c = 0; c < ncells; c++) {
        uint32_t n = cnt[c + 1];
        uint32_t room = cap_room.empty() ? kBucketCap : cap_room[c];
        uint32_t k = std::min(n, room);
        for (uint32_t j = 0; j < k; j++) kept.push_back(order[start[c] + j]);
        prefix[c + 1] = prefix[c] + k;
    }
}

// inverse(mat4) by cofactors, column-major
void invert4(const float* m, float* t) {
    float a00 = m[0], a01 = m[1], a02 = m[2], a03 = m[3];
    float a10 = m[4], a11 = m[5], a12 = m[6], a13 = m[7];
    float a20 = m[8], a21 = m[9], a22 = m[10], a23 = m[11];
    float a30 = m[12], a31 = m[13], a32 = m[14], a33 = m[15];
    float b00 = a00 * a11 - a01 * a10, b01 = a00 * a12 - a02 * a10;
    float b02 = a00 * a13 - a03 * a10, b03 = a01 * a12 - a02 * a11;
    float b04 = a01 * a13 - a03 * a11, b05 = a02 * a13 - a03 * a12;
    float b06 = a20 * a31 - a21 * a30, b07 = a20 * a32 - a22 * a30;
    float b08 = a20 * a33 - a23 * a30, b09 = a21 * a32 - a22 * a31;
    float b10 = a21 * a33 - a23 * a31, b11 = a22 * a33 - a23 * a32;
    float det = b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06;
    float id = 1.0f / det;
    t[0] = (a11 * b11 - a12 * b10 + a13 * b09) * id;
    t[1] = (a02 * b10 - a01 * b11 - a03 * b09) * id;
    t[2] = (a31 * b05 - a32 * b04 + a33 * b03) * id;
    t[3] = (a22 * b04 - a21 * b05 - a23 * b03) * id;
    t[4] = (a12 * b08 - a10 * b11 - a13 * b07) * id;
    t[5] = (a00 * b11 - a02 * b08 + a03 * b07) * id;
    t[6] = (a32 * b02 - a30 * b05 - a33 * b01) * id;
    t[7] = (a20 * b05 - a22 * b02 + a23 * b01) * id;
    t[8] = (a10 * b10 - a11 * b08 + a13 * b06) * id;
    t[9] = (a01 * b08 - a00 * b10 - a03 * b06) * id;
    t[10] = (a30 * b04 - a31 * b02 + a33 * b00) * id;
    t[11] = (a21 * b02 - a20 * b04 - a23 * b00) * id;
    t[12] = (a11 * b07 - a10 * b09 - a12 * b06) * id;
    t[13] = (a00 * b09 - a01 * b07 + a02 * b06) * id;
    t[14] = (a31 * b01 - a30 * b03 - a32 * b00) * id;
    t[15] = (a20 * b03 - a21 * b01 + a22 * b00) * id;
}

// Static reference-neighbour data of the solid path.  The reference configuration never
// changes (ref_pos' = 0, solid.rs:30, 295), so the reference buckets, the corrected-gradient
// matrix L (strain.rs:101-136) and the per-pair kernel gradients are fixed at upload.
int build_solid_lists(arp_ctx* c, const ArpParticle* particles, uint32_t n, const ArpSolidParticle* solids,
                      uint32_t ns, const ArpMaterial* mats, const std::vector<uint32_t>& kept) {
    const GridP& g = c->grid;
    const float h = c->phys.h, k = c->phys.k;
    const uint32_t dim = g.dim_n;
    std::vector<uint32_t> refcell(n, ARP_INVALID_INDEX);
    for (uint64_t cc = 0; cc < c->ncells; cc++)
        for (uint32_t q = c->rP_h[cc]; q < c->rP_h[cc + 1]; q++) refcell[kept[q]] = (uint32_t)cc;
    // (arrays in the internal solid order; refx by external id)
    const std::vector<uint32_t>& ord = c->s_ord;
    const std::vector<uint32_t>& rank = c->s_rank;
    std::vector<uint32_t> part(ns), sOff(ns + 1, 0), eOff(ns + 1, 0), sPid, ePid, eSid;
    std::vector<float> refx(4 * (size_t)ns), ref(4 * (size_t)ns), Linv(16 * (size_t)ns, 0.0f), sG, eG;
    for (uint32_t s = 0; s < ns; s++)
        for (int q = 0; q < 4; q++) refx[4 * s + q] = solids[s].ref_pos[q];
    for (uint32_t t = 0; t < ns; t++) {
        part[t] = solids[ord[t]].part_id;
        for (int q = 0; q < 4; q++) ref[4 * t + q] = refx[4 * (size_t)ord[t] + q];
    }
    auto gradw = [&](const float* r, float* out) -> float {
        float rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3];
        float diff = std::fmax(h * h - rr, 0.0f);
        float gs = -6.0f * k * diff * diff;
        for (int q = 0; q < 4; q++) out[q] = gs * r[q];
        return diff;
    };
    for (uint32_t t = 0; t < ns; t++) {
        const uint32_t s = ord[t];
        sOff[t] = (uint32_t)sPid.size();
        eOff[t] = (uint32_t)ePid.size();
        const uint32_t p = part[t];
        if (p == ARP_INVALID_INDEX || p >= n) continue;
        const uint32_t mat = particles[p].mat;
        if (!is_stiff(mats[mat])) continue;
        float L[16] = {0};
        std::vector<float> gws;
        size_t s_first = sPid.size();
        const uint32_t rc = refcell[p];
        if (rc != ARP_INVALID_INDEX) {
            const uint32_t cx = rc % g.sub[0], cy = (rc / g.sub[0]) % g.sub[1], cz = rc / (g.sub[0] * g.sub[1]);
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        int64_t x = (int64_t)cx + dx, y = (int64_t)cy + dy, z = (int64_t)cz + dz;
                        if (x < 0 || y < 0 || z < 0 || x >= g.sub[0] || y >= g.sub[1] || z >= g.sub[2]) continue;
                        uint64_t cell = (uint64_t)x + g.sub[0] * ((uint64_t)y + (uint64_t)g.sub[1] * z);
                        for (uint32_t q = c->rP_h[cell]; q < c->rP_h[cell + 1]; q++) {
                            const uint32_t p2 = kept[q];
                            const uint32_t s2 = particles[p2].solid_id;
                            const uint32_t mat2 = particles[p2].mat;
                            float dX[4], gw[4];
                            for (int r = 0; r < 4; r++) dX[r] = refx[4 * (size_t)s2 + r] - refx[4 * (size_t)s + r];
                            const float diff = gradw(dX, gw);
                            if (is_stiff(mats[mat2])) {
                                const float V2 = mats[mat2].mass / mats[mat2].start_den;
                                for (int j = 0; j < 4; j++)
                                    for (int i = 0; i < 4; i++) L[4 * j + i] += V2 * (gw[i] * dX[j]);
                                if (diff > 0.0f) {
                                    sPid.push_back(p2);
                                    for (int r = 0; r < 3; r++) gws.push_back(gw[r]);
                                    sG.push_back(0.f), sG.push_back(0.f), sG.push_back(0.f), sG.push_back(V2);
                                }
                            }
                            if (mat2 == mat && diff > 0.0f) {
                                ePid.push_back(p2);
                                eSid.push_back(rank[s2]);
                                eG.push_back(gw[0]);
                                eG.push_back(gw[1]);
                                eG.push_back(gw[2]);
                                eG.push_back(k * diff * diff * diff);
                            }
                        }
                    }
        }
        for (uint32_t i = dim; i < 4; i++) L[5 * i] = 1.0f;
        float Li[16];
        invert4(L, Li);
        for (int q = 0; q < 16; q++) Linv[16 * (size_t)t + q] = Li[q];
        // strain pairs carry L^-1 gradW(dX) (strain.rs:156)
        for (size_t q = s_first, t = 0; q < sPid.size(); q++, t++) {
            const float* gw = &gws[3 * t];
            for (int r = 0; r < 3; r++) sG[4 * q + r] = Li[r] * gw[0] + Li[4 + r] * gw[1] + Li[8 + r] * gw[2];
        }
    }
    sOff[ns] = (uint32_t)sPid.size();
    eOff[ns] = (uint32_t)ePid.size();
    // the reverse pass's weight of solid t in the F of each strain neighbour s2:
    // L_s2^-1 gradW(X_t - X_s2), static like the reference positions and L^-1 (the operations
    // of the kernel's gradient, in the same order)
    std::vector<float> sG2(sG.size(), 0.0f);
    std::vector<uint32_t> sSid(sPid.size());  // (an entry without a solid: its own, weight 0)
    for (uint32_t t = 0; t < ns; t++)
        for (uint32_t q = sOff[t]; q < sOff[t + 1]; q++) {
            const uint32_t sid2 = particles[sPid[q]].solid_id;
            sSid[q] = sid2 < ns ? rank[sid2] : t;
            if (sid2 >= ns) continue;
            const uint32_t s2 = rank[sid2];
            float r4[4], gw[4];
            for (int j = 0; j < 4; j++) r4[j] = ref[4 * (size_t)t + j] - ref[4 * (size_t)s2 + j];
            gradw(r4, gw);
            const float* L2 = &Linv[16 * (size_t)s2];
            for (int r = 0; r < 3; r++) sG2[4 * (size_t)q + r] = L2[r] * gw[0] + L2[4 + r] * gw[1] + L2[8 + r] * gw[2];
        }
    int rc;
    if ((rc = upload(c, c->part_d, part.data(), 4 * part.size()))) return rc;
    if ((rc = upload(c, c->ref_d, ref.data(), 4 * ref.size()))) return rc;
    if ((rc = upload(c, c->Linv_d, Linv.data(), 4 * Linv.size()))) return rc;
    if ((rc = upload(c, c->sOff_d, sOff.data(), 4 * sOff.size()))) return rc;
    if ((rc = upload(c, c->sPid_d, sPid.data(), 4 * sPid.size()))) return rc;
    if ((rc = upload(c, c->sG_d, sG.data(), 4 * sG.size()))) return rc;
    if ((rc = upload(c, c->sG2_d, sG2.data(), 4 * sG2.size()))) return rc;
    if ((rc = upload(c, c->sSid_d, sSid.data(), 4 * sSid.size()))) return rc;
    if ((rc = upload(c, c->eOff_d, eOff.data(), 4 * eOff.size()))) return rc;
    if ((rc = upload(c, c->ePid_d, ePid.data(), 4 * ePid.size()))) return rc;
    if ((rc = upload(c, c->eSid_d, eSid.data(), 4 * eSid.size()))) return rc;
    if ((rc = upload(c, c->eG_d, eG.data(), 4 * eG.size()))) return rc;
    return ARP_OK;
}

// ---- the sort chain ---------------------------------------------------------------
struct SortIO {
    int full;
    const float* srcP;
    const float* srcV;
    const uint32_t* srcMk;
    const uint32_t* srcPid;
    float* dstP;
    float* dstV;
    uint32_t* dstMk;
    uint32_t* dstPid;
    uint32_t* srcOut = nullptr;  // full reorder: sorted slot -> source index
    uint32_t* nlw = nullptr;  // neighbour-list words: the sort is skipped when reused
    int copyOnReuse = 0;            // full build of a reusing stage: copy src -> dst
    uint32_t* recMk = nullptr;      // full build of a recording: the recorded meta
    uint32_t* recPid = nullptr;
    uint32_t* clear = nullptr;      // words the scan zeroes first (a recording's stage flags)
    uint32_t nclear = 0;
    const float* srcAuxP = nullptr; // full build: a second state pair reordered along
    const float* srcAuxV = nullptr;
    float* dstAuxP = nullptr;
    float* dstAuxV = nullptr;
};

int run_key(arp_ctx* c, const float* posd, const uint32_t* nlw = nullptr, bool copy_back = false) {
    KeyArgs k{c->grid, c->n, posd, P<uint32_t>(c->keys), P<uint32_t>(c->ranks), P<uint32_t>(c->binCount), nlw};
    if (copy_back) {  // the key pass of a rebuild from the sorted buffers restores the landing ones
        k.cpyVelq = P<float>(c->Ysort[1]);
        k.cpyDstP = P<float>(c->Yland[0]);
        k.cpyDstV = P<float>(c->Yland[1]);
    }
    // (grid-stride: with lists kept the launch mostly returns at once, so it stays small)
    return launch(c, K_KEY, nlw ? gated_blocks(c->n) : nblk(c->n), 256, k);
}

// Yland <- Ysort when the live state is only in the sorted buffers
int sync_yland(arp_ctx* c) {
    if (!c->yland_stale) return ARP_OK;
    if (c->capture) return fail(c, ARP_E_STATE, "state copy inside a captured step");
    HIPCHK(c, hipMemcpyAsync(c->Yland[0].p, c->Ysort[0].p, 16 * (size_t)c->n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->Yland[1].p, c->Ysort[1].p, 16 * (size_t)c->n, hipMemcpyDeviceToDevice, c->stream));
    c->yland_stale = false;
    return ARP_OK;
}

int run_sort(arp_ctx* c, const SortIO& io) {
    int rc;
    const uint32_t nbins = c->grid.nrows + 1;
    const uint32_t ntiles = (nbins + kScanTile - 1) / kScanTile;
    ScanArgs s{nbins, ntiles, P<uint32_t>(c->binCount), P<uint32_t>(c->binStart),
               P<uint32_t>(c->activeRows), P<uint32_t>(c->numActive),
               P<unsigned long long>(c->scanState), P<uint32_t>(c->scanTicket), P<uint32_t>(c->scanEpoch),
               P<uint32_t>(c->stats), c->slot.stamp, io.nlw, io.clear, io.nclear};
    if ((rc = launch(c, K_SCAN, ntiles, 256, s))) return rc;
    ScatterArgs sc{};
    sc.grid = c->grid;
    sc.n = c->n;
    sc.keys = P<uint32_t>(c->keys);
    sc.ranks = P<uint32_t>(c->ranks);
    sc.binStart = P<uint32_t>(c->binStart);
    sc.tmpKey = P<uint32_t>(c->tmpKey);
    sc.tmpSrc = P<uint32_t>(c->tmpSrc);
    sc.full = io.full;
    GridSlot& g = c->slot;
    sc.perm = g.perm;
    sc.srcPosd = io.srcP;
    sc.srcVelq = io.srcV;
    sc.srcMk = io.srcMk;
    sc.srcPid = io.srcPid;
    sc.dstPosd = io.dstP;
    sc.dstVelq = io.dstV;
    sc.dstMk = io.dstMk;
    sc.dstPid = io.dstPid;
    sc.inv = io.full ? P<uint32_t>(c->inv) : nullptr;
    sc.sflag = g.sflag;
    sc.srcOut = io.full ? io.srcOut : nullptr;
    sc.nlw = io.nlw;
    sc.stamp = g.stamp;
    sc.copyOnReuse = io.copyOnReuse;
    sc.recMk = io.full ? io.recMk : nullptr;
    sc.recPid = io.full ? io.recPid : nullptr;
    if (io.full && io.dstAuxP) {
        sc.srcAuxP = io.srcAuxP;
        sc.srcAuxV = io.srcAuxV;
        sc.dstAuxP = io.dstAuxP;
        sc.dstAuxV = io.dstAuxV;
    }
    // (grid-stride: with lists kept the launch is mostly a no-op, so it stays small)
    if ((rc = launch(c, K_SCATTER, io.nlw ? gated_blocks(c->n) : nblk(c->n), 256, sc))) return rc;
    RowSortArgs r{};
    r.grid = c->grid;
    r.n = c->n;
    r.binStart = P<uint32_t>(c->binStart);
    r.activeRows = P<uint32_t>(c->activeRows);
    r.numActive = P<uint32_t>(c->numActive);
    r.tmpKey = P<uint32_t>(c->tmpKey);
    r.tmpSrc = P<uint32_t>(c->tmpSrc);
    r.gPid = P<uint32_t>(c->gPid);
    r.gGrp = P<uint32_t>(c->gGrp);
    r.rowOff = g.rowOff;
    r.rowStamp = g.rowStamp;
    r.stamp = g.stamp;
    g.full = io.full != 0;
    r.bP = P<uint32_t>(c->bP);
    r.rP = P<uint32_t>(c->rP);
    r.full = io.full;
    r.perm = g.perm;
    r.srcPosd = io.srcP;
    r.srcVelq = io.srcV;
    r.srcMk = io.srcMk;
    r.srcPid = io.srcPid;
    r.dstPosd = io.dstP;
    r.dstVelq = io.dstV;
    r.dstMk = io.dstMk;
    r.dstPid = io.dstPid;
    r.inv = io.full ? P<uint32_t>(c->inv) : nullptr;
    r.stats = P<uint32_t>(c->stats);
    r.sflag = g.sflag;
    r.srcOut = io.full ? io.srcOut : nullptr;
    r.nlw = io.nlw;
    r.recMk = io.full ? io.recMk : nullptr;
    r.recPid = io.full ? io.recPid : nullptr;
    if (io.full && io.dstAuxP) {
        r.srcAuxP = io.srcAuxP;
        r.srcAuxV = io.srcAuxV;
        r.dstAuxP = io.dstAuxP;
        r.dstAuxV = io.dstAuxV;
    }
    // one workgroup per occupied row, looping; enough workgroups for about rowsort_div
    // particles each (a tuning knob), at most 4096
    const uint32_t want = std::max<uint32_t>(1, (c->n + c->rowsort_div - 1) / c->rowsort_div);
    uint32_t blocks = std::min<uint32_t>(std::min<uint32_t>(std::max<uint32_t>(c->grid.nrows, 1), 4096), want);
    return launch(c, K_ROWSORT, blocks, 256, r);
}

FluidArgs fluid_base(arp_ctx* c) {
    FluidArgs f{};
    f.grid = c->grid;
    f.phys = c->phys;
    f.n = c->n;
    f.mats = P<DMat>(c->mats);
    f.inter = P<DInter>(c->inter);
    f.perm = c->slot.perm;
    f.rowOff = c->slot.rowOff;
    f.rowStamp = c->slot.rowStamp;
    f.stamp = c->slot.stamp;
    f.identity = c->slot.full ? 1 : 0;
    f.sflag = c->slot.sflag;
    f.stats = P<uint32_t>(c->stats);
    f.bP = P<uint32_t>(c->bP);
    f.bposd = P<float>(c->bposd);
    f.bvelq = P<float>(c->bvelq);
    f.bmk = P<uint32_t>(c->bmk);
    return f;
}

// ---- neighbour lists kept across stages --------------------------------------------
constexpr float kEps32 = 1.1920929e-7f;

// largest coordinate magnitude of the grid box
float box_extent(const GridP& g) {
    float M = 0.0f;
    for (int i = 0; i < 3; i++) M = std::max(M, std::max(std::fabs(g.min[i]), std::fabs(g.min[i] + g.dim[i])));
    return M;
}

// Lists are kept for fluid-only scenes (the solid path searches the grid itself).
bool nl_enabled(const arp_ctx* c) {
    // (the walks address list entries with 32-bit byte offsets)
    return c->nl_skin > 0.0f && !c->solids_active && c->n != 0 && !c->nl_suspend &&
           nl_fits_u32(c->nl_cap, c->n);
}

int nl_prepare(arp_ctx* c) {
    int rc;
    const size_t n = std::max<uint32_t>(c->n, 1) + kPad;
    const bool fresh = c->nlWords.p == nullptr;
    if ((rc = dalloc(c, c->nlWords, 4 * kNlWords)) || (rc = dalloc(c, c->nlList, 4 * n * c->nl_cap)) ||
        (rc = dalloc(c, c->nlCount, 4 * n)) || (rc = dalloc(c, c->nlRef, 16 * n)) ||
        (rc = dalloc(c, c->nlCell, 8 * (size_t)c->ncells)) || (rc = dalloc(c, c->nlCellRec, 8 * n)))
        return rc;
    if (c->nl_sskin > 0.0f &&
        ((rc = dalloc(c, c->nlSList, 4 * n * c->nl_cap)) || (rc = dalloc(c, c->nlSCount, 4 * n)) ||
         (rc = dalloc(c, c->nlSRef, 16 * n))))
        return rc;
    if (fresh) {
        if ((rc = fill_u32(c, c->nlWords, kNlWords, 0))) return rc;
        SpecRestoreArgs sr{P<uint32_t>(c->nlWords), nullptr, nullptr, nullptr};  // no abort pending
        if ((rc = launch(c, K_SPEC_RESTORE, 1, 64, sr))) return rc;
    }
    if (c->keys_valid) {  // the plain step's epilogue binned the state; the list step bins per stage
        if ((rc = fill_u32(c, c->binCount, c->grid.nrows + 1, 0))) return rc;
        c->keys_valid = false;
    }
    if (c->nl_dirty) {
        if ((rc = fill_u32(c, c->nlWords, 7, kInvalid))) return rc;
        FillArgs fa{P<uint32_t>(c->nlWords) + 14, 1, c->nl_max_age};
        if ((rc = launch(c, K_FILL, 1, 256, fa))) return rc;
        FillArgs fs{P<uint32_t>(c->nlWords) + 16, 2, kInvalid};  // no short list found stale
        if ((rc = launch(c, K_FILL, 1, 256, fs))) return rc;
        SpecRestoreArgs sr{P<uint32_t>(c->nlWords), nullptr, nullptr, nullptr};  // no abort pending
        if ((rc = launch(c, K_SPEC_RESTORE, 1, 64, sr))) return rc;
        if ((rc = fill_u32(c, c->nlCell, (uint32_t)(2 * c->ncells), 0))) return rc;
        c->nl_dirty = false;
    }
    return ARP_OK;
}

NlArgs nl_args(arp_ctx* c) {
    NlArgs a{};
    a.words = P<uint32_t>(c->nlWords);
    a.list = P<uint32_t>(c->nlList);
    a.count = P<uint32_t>(c->nlCount);
    a.ref = P<float>(c->nlRef);
    a.slist = P<uint32_t>(c->nlSList);
    a.scount = P<uint32_t>(c->nlSCount);
    a.sref = P<float>(c->nlSRef);
    a.cellCnt = P<uint32_t>(c->nlCell);
    a.cellRec = P<uint32_t>(c->nlCellRec);
    a.rP = P<uint32_t>(c->rP);
    a.ncells = (uint32_t)c->ncells;
    a.cap = c->nl_cap;
    if (c->nl_thresh < 0) {
        uint32_t maxStatic = 0;
        for (uint64_t q = 0; q < c->ncells; q++)
            maxStatic = std::max(maxStatic, (c->bP_h[q + 1] - c->bP_h[q]) + (c->rP_h[q + 1] - c->rP_h[q]));
        c->nl_thresh = maxStatic >= kBucketCap ? 0 : (int64_t)(kBucketCap - maxStatic);
    }
    a.capThresh = (uint32_t)c->nl_thresh;
    const float h = c->phys.h, R = h * (1.0f + c->nl_skin);
    a.rad2 = R * R;
    // the stencil covers rad with room for the rounding of x -+ rad and of the list test
    a.rad = R * (1.0f + 1e-4f) + 8.0f * kEps32 * box_extent(c->grid);
    // a particle may move (skin/2)(1 - 1e-3) from its build position: pairs within h then
    // stay within h + skin with a relative margin far above the rounding of the tests
    const float lim = 0.5f * c->nl_skin * h * (1.0f - 1e-3f);
    a.lim2 = c->nl_always ? -1.0f : lim * lim;
    a.lim = c->nl_always ? -1.0f : lim;
    // two-level lists: pairs within h stay within h + short skin while both particles stay
    // within slim of their positions at the short list's refresh (the same margin as lim)
    if (c->nl_sskin > 0.0f) {
        const float Rs = h * (1.0f + c->nl_sskin), slim = 0.5f * c->nl_sskin * h * (1.0f - 1e-3f);
        a.srad2 = Rs * Rs;
        a.slim2 = slim * slim;
    }
    // a particle within lim of its build position lies in a bucket within ext of that
    // position, with room for the rounding of the bucket function
    a.ext = lim * (1.0f + 1e-3f) + 8.0f * kEps32 * box_extent(c->grid);
    // The bucket of a coordinate is off by at most sub * 1.5 ulp of the bucket index (the
    // rounding of (p - min) / dim * sub); buckets are at least h wide up to that rounding, so
    // a pair closer than h (1 - 2 w) always lies in neighbouring buckets.
    uint32_t smax = std::max(c->grid.sub[0], std::max(c->grid.sub[1], c->grid.sub[2]));
    const float w = 4.0f * (float)smax * kEps32 + 1e-6f;
    a.grey2 = h * h * std::max(0.0f, 1.0f - 2.0f * w);
    return a;
}

// Entries per particle and stage of a recording: with kept lists at least the lists'
// capacity, so that a stage walking its list never overflows the record (the neighbours
// within h are list entries); a stage that searched can overflow, and its grid is recorded.
uint32_t rec_cap(const arp_ctx* c) { return nl_enabled(c) ? std::max<uint32_t>(kRecNb, c->nl_cap) : kRecNb; }

GridSlot ctx_slot(arp_ctx* c) {
    GridSlot g;
    g.perm = P<uint32_t>(c->perm);
    g.sflag = P<uint32_t>(c->sflag);
    g.rowOff = P<uint32_t>(c->rowOff);
    g.rowStamp = P<uint32_t>(c->rowStamp);
    g.stamp = P<uint32_t>(c->stampDev);
    return g;
}

int alloc_state(arp_ctx* c) {
    // + kPad: k_fluid reads candidate batches past the end of a range (values unused)
    size_t n = std::max<uint32_t>(c->n, 1) + kPad;
    int rc = 0;
    for (int k = 0; k < 2; k++) {
        if ((rc = dalloc(c, c->Ysort[k], 16 * n))) return rc;
        if ((rc = dalloc(c, c->Yland[k], 16 * n))) return rc;
        if ((rc = dalloc(c, c->meta_mk[k], 4 * n))) return rc;
        if ((rc = dalloc(c, c->meta_pid[k], 4 * n))) return rc;
        if ((rc = dalloc(c, c->ps_mk[k], 4 * n))) return rc;
        if ((rc = dalloc(c, c->ps_pid[k], 4 * n))) return rc;
        if ((rc = dalloc(c, c->Xb[k][0], 16 * n))) return rc;
        if ((rc = dalloc(c, c->Xb[k][1], 16 * n))) return rc;
        if ((rc = dalloc(c, c->K[0][k], 16 * n))) return rc;
    }
    if ((rc = dalloc(c, c->aos_tmp, sizeof(ArpParticle) * n))) return rc;
    if ((rc = dalloc(c, c->inv, 4 * n))) return rc;
    if ((rc = dalloc(c, c->sflag, 4 * n))) return rc;
    if (c->solids_active) {
        size_t ns = std::max<uint32_t>(c->ns, 1);
        for (int k = 0; k < 2; k++) {
            if ((rc = dalloc(c, c->KF[k], 16 * n))) return rc;
            if ((rc = dalloc(c, c->Xeps[k], 64 * ns))) return rc;
        }
        for (DBuf* b : {&c->Yeps, &c->keps, &c->ps_eps})
            if ((rc = dalloc(c, *b, 64 * ns))) return rc;
        for (DBuf* b : {&c->A_d, &c->Fm_d})  // three rows of three floats per solid
            if ((rc = dalloc(c, *b, 36 * ns))) return rc;
        // per grid row: the self-contact row test's displacement box (two float4)
        if ((rc = dalloc(c, c->rowBox, 32 * (size_t)std::max<uint32_t>(c->grid.nrows, 1)))) return rc;
    }
    for (DBuf* b : {&c->keys, &c->ranks, &c->tmpKey, &c->tmpSrc, &c->perm, &c->gPid, &c->gGrp})
        if ((rc = dalloc(c, *b, 4 * n))) return rc;
    uint32_t nb = c->grid.nrows + 1;
    if ((rc = dalloc(c, c->binCount, 4 * (size_t)nb))) return rc;
    if ((rc = dalloc(c, c->binStart, 4 * ((size_t)nb + 1)))) return rc;
    if ((rc = dalloc(c, c->activeRows, 4 * (size_t)nb))) return rc;
    if ((rc = dalloc(c, c->numActive, 16))) return rc;
    if ((rc = dalloc(c, c->stats, 16))) return rc;
    {
        const uint32_t ntiles = (nb + kScanTile - 1) / kScanTile;
        if ((rc = dalloc(c, c->scanState, 16 * (size_t)ntiles))) return rc;
        if ((rc = dalloc(c, c->scanTicket, 16))) return rc;
        if ((rc = dalloc(c, c->scanEpoch, 16))) return rc;
        HIPCHK(c, hipMemsetAsync(c->scanState.p, 0, 16 * (size_t)ntiles, c->stream));
        HIPCHK(c, hipMemsetAsync(c->scanTicket.p, 0, 16, c->stream));
        HIPCHK(c, hipMemsetAsync(c->scanEpoch.p, 0, 16, c->stream));
    }
    if ((rc = dalloc(c, c->rowOff, 4 * (size_t)c->grid.nrows * c->grid.rowstride))) return rc;
    if ((rc = dalloc(c, c->rowStamp, 4 * (size_t)c->grid.nrows))) return rc;
    HIPCHK(c, hipMemsetAsync(c->rowStamp.p, 0, 4 * (size_t)c->grid.nrows, c->stream));
    if ((rc = dalloc(c, c->stampDev, 16))) return rc;
    HIPCHK(c, hipMemsetAsync(c->stampDev.p, 0, 16, c->stream));
    c->slot = ctx_slot(c);
    c->rec_valid = false;
    return ARP_OK;
}

int ensure_k(arp_ctx* c, int slots, bool acc) {
    size_t n = std::max<uint32_t>(c->n, 1) + kPad;
    int rc;
    for (int k = 0; k < slots; k++)
        for (int j = 0; j < 2; j++)
            if ((rc = dalloc(c, c->K[k][j], 16 * n))) return rc;
    if (acc)
        for (int j = 0; j < 2; j++)
            if ((rc = dalloc(c, c->Acc[j], 16 * n))) return rc;
    if (c->solids_active) {
        size_t ns = std::max<uint32_t>(c->ns, 1);
        for (int k = 0; k < slots; k++)
            if ((rc = dalloc(c, c->Ke[k], 64 * ns))) return rc;
        if (acc && (rc = dalloc(c, c->AccE, 64 * ns))) return rc;
    }
    return ARP_OK;
}

// AoS (device) -> SoA landing buffers in original order
int aos_to_soa(arp_ctx* c, const void* d_aos, float* posd, float* velq, uint32_t* mk, uint32_t* pid) {
    AosInArgs a{c->n, d_aos, P<DMat>(c->mats), P<uint32_t>(c->mkStatic), posd, velq, mk, pid};
    return launch(c, K_AOS_IN, nblk(c->n), 256, a);
}

int soa_to_aos(arp_ctx* c, const float* posd, const float* velq, const uint32_t* pid, void* d_aos) {
    AosOutArgs a{c->n, posd, velq, pid, P<uint32_t>(c->sidExtByPid), P<uint32_t>(c->matByPid), d_aos};
    return launch(c, K_AOS_OUT, nblk(c->n), 256, a);
}

SolidArgs solid_base(arp_ctx* c) {
    SolidArgs a{};
    a.grid = c->grid;
    a.phys = c->phys;
    a.ns = c->ns;
    a.mats = P<DMat>(c->mats);
    a.part = P<uint32_t>(c->part_d);
    a.inv = P<uint32_t>(c->inv);
    a.ref = P<float>(c->ref_d);
    a.Linv = P<float>(c->Linv_d);
    a.sOff = P<uint32_t>(c->sOff_d);
    a.sPid = P<uint32_t>(c->sPid_d);
    a.sG = P<float>(c->sG_d);
    a.sG2 = P<float>(c->sG2_d);
    a.sSid = P<uint32_t>(c->sSid_d);
    a.eOff = P<uint32_t>(c->eOff_d);
    a.ePid = P<uint32_t>(c->ePid_d);
    a.eSid = P<uint32_t>(c->eSid_d);
    a.eG = P<float>(c->eG_d);
    a.A = P<float>(c->A_d);
    a.Fm = P<float>(c->Fm_d);
    a.keps = P<float>(c->keps);
    a.rowBox = P<float>(c->rowBox);
    a.perm = c->slot.perm;
    a.rowOff = c->slot.rowOff;
    a.rowStamp = c->slot.rowStamp;
    a.stamp = c->slot.stamp;
    a.sflag = c->slot.sflag;
    a.sidByPid = P<uint32_t>(c->sidByPid);
    a.KFp = P<float>(c->KF[0]);
    a.KFv = P<float>(c->KF[1]);
    a.kscale = c->kscale.p ? P<float>(c->kscale) : nullptr;
    return a;
}

// One RK step of length dt on the internal state (Yland/meta[cur] on entry and exit).
// With rec, the step also keeps what its adjoint reads: the sorted y_n and every stage
// input, the sorted meta, the reorder map and each stage's neighbour structure.
// Velocity Verlet: D = F(y) at the live state (the first step after the state was set, as
// init_with_vel), then the kick and drift of the step in place.  The evaluation sorts the
// state into the storage order it leaves in the landing buffers, so D and y share it.
int verlet_prepare(arp_ctx* c, float dt) {
    int rc;
    const size_t n = std::max<uint32_t>(c->n, 1) + kPad;
    if ((rc = dalloc(c, c->Dv[0], 16 * n)) || (rc = dalloc(c, c->Dv[1], 16 * n))) return rc;
    for (int k = 0; k < 2; k++)
        if ((rc = dalloc(c, c->XF[k][0], 16 * n)) || (rc = dalloc(c, c->XF[k][1], 16 * n))) return rc;
    if (c->solids_active && ((rc = dalloc(c, c->De, 64 * (size_t)std::max<uint32_t>(c->ns, 1))) ||
                             (rc = dalloc(c, c->XFe, 64 * (size_t)std::max<uint32_t>(c->ns, 1)))))
        return rc;
    if (!c->verlet_valid) {
        if (!c->keys_valid && (rc = run_key(c, P<float>(c->Yland[0])))) return rc;
        const int nx = c->cur ^ 1;
        const GridSlot own = c->slot;
        SortIO io{1,
                  P<float>(c->Yland[0]), P<float>(c->Yland[1]), P<uint32_t>(c->meta_mk[c->cur]),
                  P<uint32_t>(c->meta_pid[c->cur]), P<float>(c->Ysort[0]), P<float>(c->Ysort[1]),
                  P<uint32_t>(c->meta_mk[nx]), P<uint32_t>(c->meta_pid[nx])};
        if ((rc = run_sort(c, io))) return rc;
        c->cur = nx;
        c->keys_valid = false;
        c->nl_dirty = true;  // the storage order changed outside the kept lists
        FluidArgs f = fluid_base(c);
        f.Xposd = P<float>(c->Ysort[0]);
        f.Xvelq = P<float>(c->Ysort[1]);
        f.mk = P<uint32_t>(c->meta_mk[c->cur]);
        f.mode = 2;
        f.KdP = P<float>(c->Dv[0]);
        f.KdV = P<float>(c->Dv[1]);
        if ((rc = launch(c, K_FLUID, nblk(c->n), 256, f))) return rc;
        if (c->solids_active) {
            SolidArgs sa = solid_base(c);
            sa.Xposd = P<float>(c->Ysort[0]);
            sa.Xvelq = P<float>(c->Ysort[1]);
            sa.mk = P<uint32_t>(c->meta_mk[c->cur]);
            sa.pid = P<uint32_t>(c->meta_pid[c->cur]);
            sa.Xeps = P<float>(c->Yeps);
            sa.mode = 2;
            sa.KFp = P<float>(c->Dv[0]);
            sa.KFv = P<float>(c->Dv[1]);
            sa.KdP = P<float>(c->Dv[0]);
            sa.KdV = P<float>(c->Dv[1]);
            sa.keps = P<float>(c->De);
            if ((rc = launch(c, K_STRAIN, nblk(c->ns), 256, sa))) return rc;
            if ((rc = launch(c, K_SOLID, nblk(c->ns), 256, sa))) return rc;
        }
        // the state in the order of D
        HIPCHK(c, hipMemcpyAsync(c->Yland[0].p, c->Ysort[0].p, 16 * (size_t)c->n, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->Yland[1].p, c->Ysort[1].p, 16 * (size_t)c->n, hipMemcpyDeviceToDevice, c->stream));
        c->slot = own;
        c->verlet_valid = true;
    }
    if (c->keys_valid) {  // the bins counted for the old positions are void
        if ((rc = fill_u32(c, c->binCount, c->grid.nrows + 1, 0))) return rc;
        c->keys_valid = false;
    }
    VerletArgs v{c->n, c->solids_active ? c->ns : 0u, dt, P<float>(c->Yland[0]), P<float>(c->Yland[1]),
                 P<float>(c->XF[0][0]), P<float>(c->XF[0][1]),
                 P<float>(c->Dv[0]), P<float>(c->Dv[1]), P<uint32_t>(c->meta_mk[c->cur]), P<DMat>(c->mats),
                 c->solids_active ? P<float>(c->Yeps) : nullptr, c->solids_active ? P<float>(c->XFe) : nullptr,
                 c->solids_active ? P<float>(c->De) : nullptr};
    return launch(c, K_VERLET_PRE, nblk(std::max(c->n, v.ns)), 256, v);
}

// A speculative step (SpecArg): a list-reusing step issued as its walk kernels alone.
struct SpecArg {
    uint32_t step;  // index among the steps in flight (the abort word records it)
    uint32_t slot;  // its snapshot slot
    uint32_t gen;   // the list counter value before its first stage
};

int step_once(arp_ctx* c, float dt, RecStep* rec = nullptr, const SpecArg* sp = nullptr) {
    const ArpTableau& T = c->tab;
    const bool verlet = T.stages == ARP_VERLET;
    const int S = verlet ? 1 : T.stages;
    int rc;
    if (verlet && rec) return fail(c, ARP_E_STATE, "reverse mode covers the Runge-Kutta tableaus");
    if (!verlet) c->verlet_valid = false;
    // a plain list-keeping step of several stages keeps y only in the sorted buffers; every
    // other step reads the landing buffers
    const bool lazy = !verlet && S > 1 && !rec && !c->capture && nl_enabled(c);
    if (!lazy && (rc = sync_yland(c))) return rc;
    // (a recorded step speculates when the previous one left its input in place)
    if (sp && !lazy && !(rec && S > 1 && !c->capture && nl_enabled(c) && c->rec_prefilled == rec->Xp[0].p))
        return fail(c, ARP_E_STATE, "speculative step outside the list-reusing path");
    // which stage derivatives must be kept for later (non-subdiagonal) stage inputs
    int kslot[ARP_MAX_STAGES];
    int nslots = 0;
    for (int j = 0; j < S; j++) {
        kslot[j] = -1;
        for (int t = j + 2; t < S; t++)
            if (T.a[t][j] != 0.0f) kslot[j] = 0;
        if (kslot[j] == 0) kslot[j] = nslots++;
    }
    if (verlet) kslot[0] = -1;
    if (nslots > kMaxK) return fail(c, ARP_E_LIMIT, "tableau needs %d stored stages (max %d)", nslots, kMaxK);
    if (!rec) c->rec_valid = false;  // the live state leaves the recorded trajectory
    bool acc = false;
    for (int j = 0; j + 1 < S; j++)
        if (T.b[j] != 0.0f) acc = true;
    if ((rc = ensure_k(c, nslots, acc))) return rc;
    // velocity Verlet: the kick and drift before the force evaluation (moves the positions)
    if (verlet && (rc = verlet_prepare(c, dt))) return rc;
    // neighbour lists kept across stages (the plain step of a fluid-only scene)
    const bool nl = nl_enabled(c);
    if (nl) {
        if (!c->capture) {
            if ((rc = nl_prepare(c))) return rc;
        } else if (c->nl_dirty || !c->nlWords.p) {
            return fail(c, ARP_E_STATE, "neighbour lists not prepared before capture");
        }
    } else {
        c->nl_dirty = true;  // the counters and words no longer follow the state
    }
    uint32_t* nlw = nl ? P<uint32_t>(c->nlWords) : nullptr;

    // stage-1 sort: full reorder of y_n into cell-major storage (with lists kept, the key
    // pass runs at every stage that sorts, gated on the device)
    const bool stale = c->yland_stale;  // (only on the lazy path)
    if (sp) {
        // (the stage-1 sort of a reusing step does nothing: no key pass)
    } else if (nl) {
        if ((rc = run_key(c, stale ? P<float>(c->Ysort[0]) : P<float>(c->Yland[0]), nlw, stale))) return rc;
    } else if (!c->keys_valid) {
        if ((rc = run_key(c, P<float>(c->Yland[0])))) return rc;
    }
    const int nx = c->cur ^ 1;
    float* Y0p = rec ? P<float>(rec->Xp[0]) : P<float>(c->Ysort[0]);
    float* Y0v = rec ? P<float>(rec->Xv[0]) : P<float>(c->Ysort[1]);
    const GridSlot own = c->slot;
    if (rec) c->slot = rec->slot[0];
    SortIO io{1,
              P<float>(c->Yland[0]), P<float>(c->Yland[1]), P<uint32_t>(c->meta_mk[c->cur]),
              P<uint32_t>(c->meta_pid[c->cur]),
              Y0p, Y0v, P<uint32_t>(c->meta_mk[nx]),
              P<uint32_t>(c->meta_pid[nx]), rec ? P<uint32_t>(rec->src) : nullptr, nlw,
              // one-stage tableaus find y_n only in the landing buffers; so does a recording
              // unless the previous recorded step also left it in this step's stage input
              (S == 1 || (rec && c->rec_prefilled != rec->Xp[0].p)) ? 1 : 0,
              rec ? P<uint32_t>(rec->mk) : nullptr, rec ? P<uint32_t>(rec->pid) : nullptr,
              // the recording's per-stage flags start at zero
              rec ? rec->fl : nullptr, rec ? 2u * ARP_MAX_STAGES : 0u};
    if (verlet) {  // the force point is reordered along with the kick base
        io.srcAuxP = P<float>(c->XF[0][0]);
        io.srcAuxV = P<float>(c->XF[0][1]);
        io.dstAuxP = P<float>(c->XF[1][0]);
        io.dstAuxV = P<float>(c->XF[1][1]);
    }
    if (!sp && (rc = run_sort(c, io))) return rc;
    if (sp && rec) c->slot.full = true;  // (as the full reorder marks its slot)
    c->cur = nx;
    c->keys_valid = false;
    if (rec) {
        rec->slot[0] = c->slot;  // (the sort, or the speculative walk, writes the recorded meta)
    }

    const float* Xp = verlet ? P<float>(c->XF[1][0]) : Y0p;  // the first stage's input
    const float* Xv = verlet ? P<float>(c->XF[1][1]) : Y0v;
    const float* Xe = verlet ? P<float>(c->XFe) : P<float>(c->Yeps);
    if (rec && rec->sEps[0].p) {  // the step's plastic strain, the first stage's input
        HIPCHK(c, hipMemcpyAsync(rec->sEps[0].p, c->Yeps.p, 64 * (size_t)c->ns, hipMemcpyDeviceToDevice, c->stream));
        Xe = P<float>(rec->sEps[0]);
    }
    for (int s = 0; s < S; s++) {
        if (s > 0 && rec) c->slot = rec->slot[s];
        if (s > 0 && !sp) {
            SortIO pio{0, nullptr, nullptr, nullptr, P<uint32_t>(c->meta_pid[c->cur]),
                       nullptr, nullptr, nullptr, nullptr, nullptr, nlw};
            if (nl && (rc = run_key(c, Xp, nlw))) return rc;
            if ((rc = run_sort(c, pio))) return rc;
            if (rec) rec->slot[s] = c->slot;
        }
        const bool last = (s == S - 1);
        RkArgs r{};
        r.Yposd = Y0p;
        r.Yvelq = Y0v;
        r.b_cur = verlet ? 0.5f * dt : dt * T.b[s];
        r.b_pos = verlet ? 0.0f : r.b_cur;
        r.vdt = verlet ? dt : 0.0f;
        r.last = last ? 1 : 0;
        if (acc && s > 0) {
            r.AccP = P<float>(c->Acc[0]);
            r.AccV = P<float>(c->Acc[1]);
            r.AccE = c->solids_active ? P<float>(c->AccE) : nullptr;
        }
        if (acc && !last) {
            r.AccPOut = P<float>(c->Acc[0]);
            r.AccVOut = P<float>(c->Acc[1]);
            r.AccEOut = c->solids_active ? P<float>(c->AccE) : nullptr;
        }
        if (!last) {
            r.a_next = dt * T.a[s + 1][s];
            r.nk = 0;
            for (int j = 0; j < s; j++)
                if (T.a[s + 1][j] != 0.0f) {
                    r.Kp[r.nk] = P<float>(c->K[kslot[j]][0]);
                    r.Kv[r.nk] = P<float>(c->K[kslot[j]][1]);
                    r.Ke[r.nk] = c->solids_active ? P<float>(c->Ke[kslot[j]]) : nullptr;
                    r.kc[r.nk] = dt * T.a[s + 1][j];
                    r.nk++;
                }
            r.Oposd = rec ? P<float>(rec->Xp[s + 1]) : P<float>(c->Xb[s % 2][0]);
            r.Ovelq = rec ? P<float>(rec->Xv[s + 1]) : P<float>(c->Xb[s % 2][1]);
            r.Oeps = (rec && rec->sEps[s + 1].p) ? P<float>(rec->sEps[s + 1]) : P<float>(c->Xeps[s % 2]);
        } else {
            r.Oposd = P<float>(c->Yland[0]);
            r.Ovelq = P<float>(c->Yland[1]);
            r.Oeps = P<float>(c->Yeps);  // in place: each solid's thread reads then writes
            // with lists kept, y_{n+1} goes to the sorted buffers, where a next step that
            // reuses its lists (and so keeps the storage order) finds it; a thread overwrites
            // only its own y_n there, and the stage's neighbours come from its own input.
            // The landing buffers are brought up to date when something reads them.
            if (lazy) {
                r.Oposd = Y0p;
                r.Ovelq = Y0v;
            } else if (nl && S > 1 && !rec) {
                r.Oposd2 = Y0p;
                r.Ovelq2 = Y0v;
            } else if (nl && S > 1 && rec && c->rec_next && c->rec_next != rec) {  // the next recorded step's input
                r.Oposd2 = P<float>(c->rec_next->Xp[0]);
                r.Ovelq2 = P<float>(c->rec_next->Xv[0]);
            }
        }
        r.Yeps = P<float>(c->Yeps);
        if (kslot[s] >= 0) {
            r.KpOut = P<float>(c->K[kslot[s]][0]);
            r.KvOut = P<float>(c->K[kslot[s]][1]);
            r.KeOut = c->solids_active ? P<float>(c->Ke[kslot[s]]) : nullptr;
        }
        if (verlet) {  // the force evaluation is kept for the next step's first kick
            r.KpOut = P<float>(c->Dv[0]);
            r.KvOut = P<float>(c->Dv[1]);
            r.KeOut = c->solids_active ? P<float>(c->De) : nullptr;
        }
        if (!nl && !verlet) {  // the next stage's keys (with lists kept: the gated key pass)
            r.keys = P<uint32_t>(c->keys);
            r.ranks = P<uint32_t>(c->ranks);
            r.binCount = P<uint32_t>(c->binCount);
        }

        FluidArgs f = fluid_base(c);
        f.Xposd = Xp;
        f.Xvelq = Xv;
        f.mk = P<uint32_t>(c->meta_mk[c->cur]);
        f.mode = 0;
        f.skip_elastic = c->solids_active ? 1 : 0;
        f.KdP = P<float>(c->KF[0]);
        f.KdV = P<float>(c->KF[1]);
        f.rk = r;
        if (rec) {  // keep the stage's neighbour lists for the reverse sweep
            f.nbList = P<uint32_t>(rec->nbList[s]);
            f.nbCount = P<uint32_t>(rec->nbCount[s]);
            f.nbOver = rec->fl + s;
            f.nbCap = c->rec_nb_cap;  // the capacity the record was allocated with
        }
        if (nl) {
            // the list pair: one searches and builds, the other walks the kept lists; the
            // device decides which of the two works (the other returns at once)
            f.nl = nl_args(c);
            if (s == 0) {  // a build after the full reorder mirrors the sorted meta
                f.mkMirror = P<uint32_t>(c->meta_mk[c->cur ^ 1]);
                f.pidMirror = P<uint32_t>(c->meta_pid[c->cur ^ 1]);
                f.pidIn = P<uint32_t>(c->meta_pid[c->cur]);
            }
            if (rec) f.recUse = rec->fl + ARP_MAX_STAGES + s;
            // one mobile and one boundary material, nothing elastic: the walk selects a
            // pair's materials by kind (a batch when every scene of it can)
            const bool fm = (!c->capture || c->capture_fm) && c->mobile_mat >= 0 && !c->any_elastic &&
                            !c->bnd_elastic && c->bnd_mat != -2;
            f.fastMat = fm ? (uint32_t)c->mobile_mat : 0u;
            f.fastBnd = fm ? (uint32_t)(c->bnd_mat >= 0 ? c->bnd_mat : c->mobile_mat) : 0u;
            if (sp) {  // the walk alone, keeping the chain's bookkeeping
                f.spec = 1;
                f.specGen = sp->gen + (uint32_t)s + 1u;
                f.specStep = (sp->step << 4) | (uint32_t)s;
                f.specSnap = s == 0 ? P<uint32_t>(c->nlSnap) + (size_t)kSpecSnap * sp->slot : nullptr;
                f.hostAbort = c->hAbortDev;
                f.stampW = c->slot.stamp;
                f.statsW = P<uint32_t>(c->stats);
                if (rec && s == 0) {
                    f.specClear = rec->fl;
                    f.recSrc = P<uint32_t>(rec->src);
                    f.recMk = P<uint32_t>(rec->mk);
                    f.recPid = P<uint32_t>(rec->pid);
                }
            } else if ((rc = launch(c, rec ? K_FLUID_NLBR : K_FLUID_NLB,
                                    // (in a batch: gated_blocks workgroups, each building
                                    // several blocks of particles)
                                    c->capture && !rec ? gated_blocks(c->n) : nblk(c->n), 256, f))) {
                return rc;
            }
            const Kern walk = rec ? (fm ? K_FLUID_NLR_F : K_FLUID_NLR) : (fm ? K_FLUID_NL_F : K_FLUID_NL);
            if ((rc = launch(c, walk, nblk(c->n), 256, f))) return rc;
        } else if ((rc = launch(c, rec ? K_FLUID_REC : K_FLUID, nblk(c->n), 256, f))) {
            return rc;
        }
        if (c->solids_active) {
            SolidArgs sa = solid_base(c);
            sa.Xposd = Xp;
            sa.Xvelq = Xv;
            sa.mk = P<uint32_t>(c->meta_mk[c->cur]);
            sa.pid = P<uint32_t>(c->meta_pid[c->cur]);
            sa.Xeps = Xe;
            // solids of non-plastic materials keep y_n's plastic strain (zero rate)
            sa.epsY = verlet ? nullptr : P<float>(c->Yeps);
            if (rec) {  // the stage's strain results stay in the record
                sa.A = P<float>(rec->sA[s]);
                sa.Fm = P<float>(rec->sF[s]);
                sa.Fd = P<float>(rec->sFd[s]);
                sa.rowBox = P<float>(rec->sBox[s]);
            }
            sa.mode = 0;
            sa.rk = r;
            if ((rc = launch(c, K_STRAIN, nblk(c->ns), 256, sa))) return rc;
            if ((rc = launch(c, K_SOLID, nblk(c->ns), 256, sa))) return rc;
        } else if (c->capture && c->capture_solid) {
            // a fluid scene in a batch with solid scenes: the solid path's launches with no
            // workgroups of this scene, so that every scene has the same launch sequence
            const SolidArgs sa{};
            if ((rc = launch(c, K_STRAIN, 0, 256, sa))) return rc;
            if ((rc = launch(c, K_SOLID, 0, 256, sa))) return rc;
        }
        if (!last) {
            Xp = r.Oposd;
            Xv = r.Ovelq;
            Xe = r.Oeps;
        }
    }
    if (rec) c->slot = own;
    c->yland_stale = lazy;
    c->keys_valid = !nl && !verlet;  // the plain step's last epilogue binned y_{n+1}
    c->rec_prefilled = (rec && nl && S > 1 && c->rec_next && c->rec_next != rec) ? c->rec_next->Xp[0].p : nullptr;
    c->have_build = false;
    return ARP_OK;
}

int need_scene(arp_ctx* c) {
    if (!c) return ARP_E_ARG;
    if (!c->have_scene) return fail(c, ARP_E_STATE, "no scene uploaded");
    return ARP_OK;
}

// ---- reverse mode ---------------------------------------------------------------------
int d2d(arp_ctx* c, DBuf& dst, const DBuf& src, size_t bytes) {
    int rc = dalloc(c, dst, bytes);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(dst.p, src.p, bytes, hipMemcpyDeviceToDevice, c->stream));
    return ARP_OK;
}

int alloc_rec(arp_ctx* c, RecStep& R, int S) {
    const size_t n = std::max<uint32_t>(c->n, 1) + kPad;
    const size_t nr = std::max<uint32_t>(c->grid.nrows, 1);
    int rc;
    for (DBuf* b : {&R.mk, &R.pid, &R.src})
        if ((rc = dalloc(c, *b, 4 * n))) return rc;
    // the slots' build counters and row stamps start at zero when (re)allocated; kept
    // buffers keep growing counters, so rows left from an older build never match
    const void* st0 = R.stamps.p;
    if ((rc = dalloc(c, R.stamps, 4 * ARP_MAX_STAGES))) return rc;
    if (R.stamps.p != st0) HIPCHK(c, hipMemsetAsync(R.stamps.p, 0, 4 * ARP_MAX_STAGES, c->stream));
    for (int s = 0; s < S; s++)
        if ((rc = dalloc(c, R.nbList[s], 4 * (size_t)c->rec_nb_cap * n)) || (rc = dalloc(c, R.nbCount[s], 4 * n))) return rc;
    for (int s = 0; s < S; s++) {
        const void* rs0 = R.rowStamp[s].p;
        if ((rc = dalloc(c, R.Xp[s], 16 * n)) || (rc = dalloc(c, R.Xv[s], 16 * n)) ||
            (rc = dalloc(c, R.perm[s], 4 * n)) || (rc = dalloc(c, R.sflag[s], 4 * n)) ||
            (rc = dalloc(c, R.rowOff[s], 4 * nr * c->grid.rowstride)) || (rc = dalloc(c, R.rowStamp[s], 4 * nr)))
            return rc;
        if (R.rowStamp[s].p != rs0 || R.stamps.p != st0)
            HIPCHK(c, hipMemsetAsync(R.rowStamp[s].p, 0, 4 * nr, c->stream));
        GridSlot& g = R.slot[s];
        g.perm = P<uint32_t>(R.perm[s]);
        g.sflag = P<uint32_t>(R.sflag[s]);
        g.rowOff = P<uint32_t>(R.rowOff[s]);
        g.rowStamp = P<uint32_t>(R.rowStamp[s]);
        g.stamp = P<uint32_t>(R.stamps) + s;
        g.full = false;
        if (c->solids_active) {
            const size_t ns = std::max<uint32_t>(c->ns, 1);
            if ((rc = dalloc(c, R.sA[s], 36 * ns)) || (rc = dalloc(c, R.sF[s], 36 * ns)) ||
                (rc = dalloc(c, R.sFd[s], 36 * ns)) || (rc = dalloc(c, R.sBox[s], 32 * nr)))
                return rc;
            if (any_plastic(c) && (rc = dalloc(c, R.sEps[s], 64 * ns))) return rc;
        }
    }
    return ARP_OK;
}

void free_buf(DBuf& b) {
    if (b.p) hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

// Keep the first nrec recorded steps and nck checkpoints (their buffers are reused by the
// next recording), free the rest.
void trim_rec(arp_ctx* c, size_t nrec, size_t nck) {
    for (size_t i = nrec; i < c->rec.size(); i++) {
        RecStep& R = c->rec[i];
        for (DBuf* b : {&R.mk, &R.pid, &R.src, &R.stamps}) free_buf(*b);
        for (int s = 0; s < ARP_MAX_STAGES; s++) free_buf(R.nbList[s]), free_buf(R.nbCount[s]);
        for (int s = 0; s < ARP_MAX_STAGES; s++)
            for (DBuf* b : {&R.Xp[s], &R.Xv[s], &R.perm[s], &R.sflag[s], &R.rowOff[s], &R.rowStamp[s], &R.sA[s],
                            &R.sF[s], &R.sFd[s], &R.sBox[s], &R.sEps[s]})
                free_buf(*b);
    }
    for (size_t i = nck; i < c->ck.size(); i++)
        for (DBuf* b : {&c->ck[i].P, &c->ck[i].V, &c->ck[i].mk, &c->ck[i].pid, &c->ck[i].E}) free_buf(*b);
    if (c->rec.size() > nrec) c->rec.resize(nrec);
    if (c->ck.size() > nck) c->ck.resize(nck);
    c->rec_valid = false;
}

void free_rec(arp_ctx* c) { trim_rec(c, 0, 0); }

// the live state (y in its storage order and the meta of that order); without `full` only the
// order (a recording of one segment is never replayed: its checkpoints give only orders)
int ck_save(arp_ctx* c, Checkpoint& k, bool full = true) {
    const size_t n = c->n;
    int rc;
    if (full && (rc = sync_yland(c))) return rc;
    if ((rc = d2d(c, k.pid, c->meta_pid[c->cur], 4 * n))) return rc;
    if (full && ((rc = d2d(c, k.P, c->Yland[0], 16 * n)) || (rc = d2d(c, k.V, c->Yland[1], 16 * n)) ||
                 (rc = d2d(c, k.mk, c->meta_mk[c->cur], 4 * n))))
        return rc;
    if (full && c->solids_active && any_plastic(c) && (rc = d2d(c, k.E, c->Yeps, 64 * (size_t)c->ns))) return rc;
    return ARP_OK;
}

// Restore a saved state.  The next step bins it again from its positions; the step is then
// the same computation as from the original state (the cell-major order is canonical: row,
// cell, particle id), so replays reproduce the recorded trajectory bit for bit.
int ck_restore(arp_ctx* c, Checkpoint& k) {
    const size_t n = c->n;
    c->nl_dirty = true;
    c->verlet_valid = false;
    c->rec_prefilled = nullptr;
    c->yland_stale = false;
    HIPCHK(c, hipMemcpyAsync(c->Yland[0].p, k.P.p, 16 * n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->Yland[1].p, k.V.p, 16 * n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->meta_mk[c->cur].p, k.mk.p, 4 * n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->meta_pid[c->cur].p, k.pid.p, 4 * n, hipMemcpyDeviceToDevice, c->stream));
    if (k.E.p && c->solids_active && any_plastic(c))
        HIPCHK(c, hipMemcpyAsync(c->Yeps.p, k.E.p, 64 * (size_t)c->ns, hipMemcpyDeviceToDevice, c->stream));
    int rc = fill_u32(c, c->binCount, c->grid.nrows + 1, 0);
    if (rc) return rc;
    c->keys_valid = false;
    c->have_build = false;
    return ARP_OK;
}

// host copies of the neighbour-record overflow flags of the first nrec recorded steps
int fetch_over(arp_ctx* c, uint32_t nrec) {
    constexpr size_t W = 2 * ARP_MAX_STAGES;
    std::vector<uint32_t> h(W * (size_t)std::max<uint32_t>(nrec, 1));
    if (nrec) HIPCHK(c, hipMemcpyAsync(h.data(), c->recFlags.p, 4 * W * nrec, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (uint32_t j = 0; j < nrec; j++) std::memcpy(c->rec[j].over, &h[W * j], sizeof(c->rec[j].over));
    return ARP_OK;
}

// Reverse sweep over one recorded step: lam (P,V) holds dL/dy_{n+1} in the step's sorted
// order on entry and dL/dy_n in the order of the state before the step on exit.
//   w_s = dt b_s lam + sum_{t>s} dt a_ts G_t,   G_s = J_f(X_s)^T w_s,   dL/dy_n = lam + sum_s G_s
int reverse_step(arp_ctx* c, RecStep& R, float dt) {
    const ArpTableau& T = c->tab;
    const int S = T.stages;
    const size_t n = c->n;
    const uint32_t nm = c->nm;
    int rc;
    (void)n;
    if (c->solids_active) {  // particle id -> storage index of this step
        InvArgs ia{c->n, P<uint32_t>(R.pid), P<uint32_t>(c->inv)};
        if ((rc = launch(c, K_INV, nblk(c->n), 256, ia))) return rc;
    }
    bool wReady = false;   // Wb already holds this stage's W (formed by the previous kernel)
    bool folded = false;   // the last kernel wrote dL/dy_n in the previous order (LamNext)
    for (int s = S - 1; s >= 0; s--) {
        AdjWArgs w{};
        w.n = c->n;
        w.cl = dt * T.b[s];
        w.LP = P<float>(c->LamIn[0]);
        w.LV = P<float>(c->LamIn[1]);
        for (int t = s + 1; t < S; t++)
            if (T.a[t][s] != 0.0f) {
                w.GP[w.ng] = P<float>(c->Gb[t][0]);
                w.GV[w.ng] = P<float>(c->Gb[t][1]);
                w.cg[w.ng] = dt * T.a[t][s];
                w.ng++;
            }
        w.WP = P<float>(c->Wb[0]);
        w.WV = P<float>(c->Wb[1]);
        // a single scaled term is read in place by the stage kernel; otherwise it is formed
        const float* wP = P<float>(c->Wb[0]);
        const float* wV = P<float>(c->Wb[1]);
        float wscale = 1.0f;
        bool wIsWb = true;
        if (wReady) {
        } else if (w.ng == 0) {
            wP = w.LP;
            wV = w.LV;
            wscale = w.cl;
            wIsWb = false;
        } else if (w.ng == 1 && w.cl == 0.0f) {
            wP = w.GP[0];
            wV = w.GV[0];
            wscale = w.cg[0];
            wIsWb = false;
        } else if ((rc = launch(c, K_ADJ_W, nblk(c->n), 256, w))) {
            return rc;
        }
        wReady = false;
        const GridSlot& g = R.slot[s];
        AdjArgs a{};
        a.grid = c->grid;
        a.phys = c->phys;
        a.n = c->n;
        a.mats = P<DMat>(c->mats);
        a.inter = P<DInter>(c->inter);
        a.perm = g.perm;
        a.rowOff = g.rowOff;
        a.rowStamp = g.rowStamp;
        a.stamp = g.stamp;
        a.identity = g.full ? 1 : 0;
        a.sflag = g.sflag;
        a.bP = P<uint32_t>(c->bP);
        a.bposd = P<float>(c->bposd);
        a.bvelq = P<float>(c->bvelq);
        a.bmk = P<uint32_t>(c->bmk);
        a.Xposd = P<float>(R.Xp[s]);
        a.Xvelq = P<float>(R.Xv[s]);
        a.mk = P<uint32_t>(R.mk);
        a.Wp = wP;
        a.Wv = wV;
        a.wscale = wscale;
        // the next reverse stage's W formed here when it is wl * Lam_in + wg * G_s alone
        bool fuseW = false;
        if (s == S - 1 && s >= 1 && !wIsWb && !c->solids_active) {
            fuseW = true;
            for (int t = s; t < S; t++)  // terms of W_{s-1}: t > s-1 with a[t][s-1] != 0
                if (T.a[t][s - 1] != 0.0f && t != s) fuseW = false;
            if (T.a[s][s - 1] == 0.0f) fuseW = false;
        }
        // G_s is kept when a later reverse stage reads it (beyond a fused W)
        bool needG = c->solids_active;
        for (int t = 0; t < s; t++)
            if (T.a[s][t] != 0.0f && !(fuseW && t == s - 1)) needG = true;
        a.Gp = needG ? P<float>(c->Gb[s][0]) : nullptr;
        a.Gv = needG ? P<float>(c->Gb[s][1]) : nullptr;
        if (fuseW) {
            a.WnP = P<float>(c->Wb[0]);
            a.WnV = P<float>(c->Wb[1]);
            a.wl = dt * T.b[s - 1];
            a.wg = dt * T.a[s][s - 1];
            wReady = true;
        }
        if (s == 0 && !c->solids_active) {  // the step's permutation, folded in
            a.permOut = P<uint32_t>(R.src);
            a.LamOutP = P<float>(c->LamNext[0]);
            a.LamOutV = P<float>(c->LamNext[1]);
            folded = true;
        }
        a.LamP = P<float>(c->LamAcc[0]);
        a.LamV = P<float>(c->LamAcc[1]);
        a.pgrad = P<float>(c->pgradPart);
        a.nbList = P<uint32_t>(R.nbList[s]);
        a.nbCount = P<uint32_t>(R.nbCount[s]);
        a.nbOver = R.fl + s;
        a.stageUse = R.fl + ARP_MAX_STAGES + s;
        if (s == S - 1) {  // the step's accumulator starts from dL/dy_{n+1}
            a.LamInP = P<float>(c->LamIn[0]);
            a.LamInV = P<float>(c->LamIn[1]);
        }
        // the recorded neighbours, unless a particle had more than the record holds
        // the fast kernels: LDS tables, one mobile material
        const bool fast = c->mobile_mat >= 0 && c->nm <= (uint32_t)kMatLds;
        a.fastMat = fast ? (uint32_t)c->mobile_mat : 0u;
        const bool fl = fast && !c->any_elastic;
        const Kern ka = R.over[s] ? (fl ? K_FLUID_ADJ_F : fast ? K_FLUID_ADJ : K_FLUID_ADJ_G)
                                  : (fl ? K_FLUID_ADJ_REC_F : fast ? K_FLUID_ADJ_REC : K_FLUID_ADJ_REC_G);
        if ((rc = launch(c, ka, nblk(c->n), 256, a))) return rc;
        if (c->solids_active) {
            // the two solid passes
            SolidArgs sa = solid_base(c);
            sa.perm = g.perm;
            sa.rowOff = g.rowOff;
            sa.rowStamp = g.rowStamp;
            sa.stamp = g.stamp;
            sa.sflag = g.sflag;
            sa.Xposd = P<float>(R.Xp[s]);
            sa.Xvelq = P<float>(R.Xv[s]);
            sa.mk = P<uint32_t>(R.mk);
            sa.pid = P<uint32_t>(R.pid);
            // the stage's input plastic strain (constant, y_n's, without plastic flow)
            sa.Xeps = R.sEps[s].p ? P<float>(R.sEps[s]) : P<float>(c->Yeps);
            // the stage's A, F and row boxes as its recorded forward pass left them
            sa.A = P<float>(R.sA[s]);
            sa.Fm = P<float>(R.sF[s]);
            sa.Fd = P<float>(R.sFd[s]);
            sa.rowBox = P<float>(R.sBox[s]);
            SolidAdjArgs sj{};
            sj.f = sa;
            sj.Wp = wP;
            sj.Wv = wV;
            sj.wscale = wscale;
            sj.Gp = P<float>(c->Gb[s][0]);
            sj.Gv = P<float>(c->Gb[s][1]);
            sj.LamP = P<float>(c->LamAcc[0]);
            sj.LamV = P<float>(c->LamAcc[1]);
            sj.Fbar = P<float>(c->Fbar);
            sj.spar = P<float>(c->spar);
            const bool other = any_other_order(c);
            sj.Ebuf = other ? P<float>(c->Ebuf) : nullptr;
            if (c->LamE.p && any_plastic(c)) {
                sj.lamE = P<float>(c->LamE);
                for (int t = 0; t < S; t++) {
                    sj.gE[t] = P<float>(c->GbE[t]);
                    sj.wcg[t] = dt * T.a[t][s];
                }
                sj.wcl = dt * T.b[s];
                sj.stage = s;
                sj.nstages = S;
            }
            if ((rc = launch(c, K_SOLID_ADJ1, nblk(c->ns), 256, sj))) return rc;
            if (other && (rc = launch(c, K_SOLID_ADJ_STRAIN, nblk(c->ns), 256, sj))) return rc;
            if ((rc = launch(c, K_SOLID_ADJ2, nblk(c->ns), 256, sj))) return rc;
        }
    }
    if (folded) {  // dL/dy_n is in LamNext
        std::swap(c->LamIn[0], c->LamNext[0]);
        std::swap(c->LamIn[1], c->LamNext[1]);
        return ARP_OK;
    }
    AdjPermArgs pa{c->n, P<uint32_t>(R.src), P<float>(c->LamAcc[0]), P<float>(c->LamAcc[1]),
                   P<float>(c->LamIn[0]), P<float>(c->LamIn[1])};
    return launch(c, K_ADJ_PERM, nblk(c->n), 256, pa);
}

// Per-stage path: device AoS state -> sorted SoA in the ps buffers (full reorder).
int stage_sort_aos(arp_ctx* c, const void* d_state) {
    int rc;
    c->nl_dirty = true;  // this build advances the grid's counter outside the step
    if ((rc = fill_u32(c, c->binCount, c->grid.nrows + 1, 0))) return rc;
    float* lp = P<float>(c->Xb[0][0]);
    float* lv = P<float>(c->Xb[0][1]);
    if ((rc = aos_to_soa(c, d_state, lp, lv, P<uint32_t>(c->ps_mk[0]), P<uint32_t>(c->ps_pid[0])))) return rc;
    if ((rc = run_key(c, lp))) return rc;
    SortIO io{1, lp, lv, P<uint32_t>(c->ps_mk[0]), P<uint32_t>(c->ps_pid[0]),
              P<float>(c->Xb[1][0]), P<float>(c->Xb[1][1]), P<uint32_t>(c->ps_mk[1]),
              P<uint32_t>(c->ps_pid[1])};
    if ((rc = run_sort(c, io))) return rc;
    c->keys_valid = false;
    c->last_pid = &c->ps_pid[1];
    c->last_posd = &c->Xb[1][0];
    c->have_build = true;
    return ARP_OK;
}

}  // namespace

// ====================================================================================
// C ABI
// ====================================================================================
extern "C" {

int arp_create(int device, uint32_t flags, arp_ctx** out) {
    (void)flags;
    if (!out) return ARP_E_ARG;
    *out = nullptr;
    arp_ctx* c = new (std::nothrow) arp_ctx();
    if (!c) return ARP_E_NOMEM;
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreate(&c->own_stream);
    // ARP_CODE_OBJECT: load a separately built gfx950 code object of the same kernels
    // (tuning variants, e.g. another ARP_BATCH) instead of the embedded one
    std::vector<char> co;
    if (const char* path = getenv("ARP_CODE_OBJECT")) {
        FILE* f = fopen(path, "rb");
        if (f) {
            fseek(f, 0, SEEK_END);
            long sz = ftell(f);
            fseek(f, 0, SEEK_SET);
            co.resize(sz > 0 ? (size_t)sz : 0);
            if (sz > 0 && fread(co.data(), 1, (size_t)sz, f) != (size_t)sz) co.clear();
            fclose(f);
        }
        if (co.empty()) {
            fprintf(stderr, "arp_create: cannot read ARP_CODE_OBJECT=%s\n", path);
            delete c;
            return ARP_E_ARG;
        }
    }
    if (e == hipSuccess) e = hipModuleLoadData(&c->mod, co.empty() ? (const void*)arp_kernels_co : (const void*)co.data());
    for (int k = 0; e == hipSuccess && k < K_COUNT; k++) e = hipModuleGetFunction(&c->fn[k], c->mod, kKernNames[k]);
    if (e != hipSuccess) {
        fprintf(stderr, "arp_create: %s\n", hipGetErrorString(e));
        if (c->own_stream) hipStreamDestroy(c->own_stream);
        delete c;
        return ARP_E_HIP;
    }
    c->stream = c->own_stream;
    // with the device current: this GPU's scratch array (looked up once per device)
    c->scratch_err = arp_scratch_acquire(device, &c->scratch, &c->scratch_bytes);
    arp_tableau_by_name("rk4", &c->tab);
    *out = c;
    return ARP_OK;
}

void arp_destroy(arp_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (DBuf* b : c->bufs)
        if (b->p) hipFree(b->p);
    free_rec(c);
    for (DBuf* b : {&c->LamIn[0], &c->LamIn[1], &c->LamAcc[0], &c->LamAcc[1], &c->LamNext[0], &c->LamNext[1],
                    &c->Wb[0], &c->Wb[1], &c->pgradPart, &c->pgradSum, &c->Fbar, &c->spar, &c->Ebuf, &c->LamE, &c->kscale, &c->gaos, &c->recFlags, &c->nlSnap})
        free_buf(*b);
    if (c->hAbort) hipHostFree(c->hAbort);
    for (hipEvent_t e : c->specEv)
        if (e) hipEventDestroy(e);
    for (int s = 0; s < ARP_MAX_STAGES; s++) free_buf(c->Gb[s][0]), free_buf(c->Gb[s][1]), free_buf(c->GbE[s]);
    for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
    if (c->mod) hipModuleUnload(c->mod);
    if (c->own_stream) hipStreamDestroy(c->own_stream);
    delete c;
}

const char* arp_last_error(const arp_ctx* c) { return c ? c->err.c_str() : "null context"; }

int arp_set_stream(arp_ctx* c, void* s) {
    if (!c) return ARP_E_ARG;
    c->stream = s ? reinterpret_cast<hipStream_t>(s) : c->own_stream;
    return ARP_OK;
}

int arp_synchronize(arp_ctx* c) {
    if (!c) return ARP_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ARP_OK;
}

float arp_norm_const(uint32_t dim, float h) {
    const float pi = 3.14159265358979323846f;
    return (pi * pi * pi) / (6.0f * vol_nball(dim + 6, h));
}

void arp_default_interaction(const ArpMaterial* a, const ArpMaterial* b, float h, ArpMatInteraction* out) {
    if (a->state_eq != 0 && b->state_eq != 0) {
        *out = ArpMatInteraction{0, 0.0f, 0.0f, 0.0f, 0.0f};
    } else {
        *out = ArpMatInteraction{2, 10000.0f, h * 2.0f, 0.0f, a->visc + b->visc};
    }
}

void arp_resolve_interactions(uint32_t nm, const ArpMaterial* mats, const ArpMatInteraction* cfg,
                              const uint8_t* present, float h, ArpMatInteraction* out) {
    for (uint32_t i = 0; i < nm; i++)
        for (uint32_t j = i; j < nm; j++) {
            ArpMatInteraction it;
            if (present && present[i * nm + j])
                it = cfg[i * nm + j];
            else if (present && present[j * nm + i])
                it = cfg[j * nm + i];
            else if (i != j)
                arp_default_interaction(&mats[i], &mats[j], h, &it);
            else
                it = ArpMatInteraction{0, 0.0f, 0.0f, 0.0f, 0.0f};
            out[i * nm + j] = it;
            out[j * nm + i] = it;
        }
}

int arp_tableau_by_name(const char* name, ArpTableau* out) {
    if (!name || !out) return ARP_E_ARG;
    std::string s(name);
    for (char& ch : s) ch = (char)tolower((unsigned char)ch);
    ArpTableau t{};
    auto set = [&](int S, std::initializer_list<float> a, std::initializer_list<float> b) {
        t.stages = S;
        int k = 0;
        for (float v : a) {
            t.a[k / S][k % S] = v;
            k++;
        }
        k = 0;
        for (float v : b) t.b[k++] = v;
        for (int i = 0; i < S; i++) {
            float cs = 0;
            for (int j = 0; j < S; j++) cs += t.a[i][j];
            t.c[i] = cs;
        }
    };
    if (s == "verlet")
        t.stages = ARP_VERLET;
    else if (s == "euler" || s == "rk1")
        set(1, {0.0f}, {1.0f});
    else if (s == "midpoint" || s == "rk2")
        set(2, {0, 0, 0.5f, 0}, {0.0f, 1.0f});
    else if (s == "heun2" || s == "heun")
        set(2, {0, 0, 1.0f, 0}, {0.5f, 0.5f});
    else if (s == "ralston")
        set(2, {0, 0, 2.0f / 3.0f, 0}, {0.25f, 0.75f});
    else if (s == "rk3")
        set(3, {0, 0, 0, 0.5f, 0, 0, -1.0f, 2.0f, 0}, {1.0f / 6.0f, 2.0f / 3.0f, 1.0f / 6.0f});
    else if (s == "heun3")
        set(3, {0, 0, 0, 1.0f / 3.0f, 0, 0, 0, 2.0f / 3.0f, 0}, {0.25f, 0.0f, 0.75f});
    else if (s == "rk4")
        set(4, {0, 0, 0, 0, 0.5f, 0, 0, 0, 0, 0.5f, 0, 0, 0, 0, 1.0f, 0},
            {1.0f / 6.0f, 1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 6.0f});
    else if (s == "rk_3_8")
        set(4, {0, 0, 0, 0, 1.0f / 3.0f, 0, 0, 0, -1.0f / 3.0f, 1.0f, 0, 0, 1.0f, -1.0f, 1.0f, 0},
            {0.125f, 0.375f, 0.375f, 0.125f});
    else
        return ARP_E_ARG;
    *out = t;
    return ARP_OK;
}

int arp_scene_upload(arp_ctx* c, const ArpParticle* particles, uint32_t n, const ArpSolidParticle* solids,
                     uint32_t ns, const ArpParticle* boundary, uint32_t nb, const ArpMaterial* mats,
                     uint32_t nm, const ArpMatInteraction* inter, const ArpAABB* bounds, float h,
                     float gravity, float artificial_viscosity) {
    (void)artificial_viscosity;  // uniform f of the reference shaders is never read (fluid.rs:43)
    if (!c || !bounds || !mats || !inter || nm == 0 || (n && !particles) || (nb && !boundary))
        return fail(c, ARP_E_ARG, "arp_scene_upload: missing argument");
    if (!(h > 0.0f)) return fail(c, ARP_E_ARG, "kernel radius must be > 0");
    if (nm > 0xFFFF) return fail(c, ARP_E_LIMIT, "at most 65535 materials");
    if (bounds->dim[3] > 0.0f) return fail(c, ARP_E_ARG, "4-D scenes (bounds.dim[3] > 0) are not supported");
    HIPCHK(c, hipSetDevice(c->device));
    GridP g{};
    uint32_t sub4[4];
    for (int i = 0; i < 4; i++) {
        float q = bounds->dim[i] / h;
        float m = 1.0f > q ? 1.0f : q;
        if (!(m == m)) m = 1.0f;
        sub4[i] = m >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)m;
    }
    for (int i = 0; i < 3; i++) {
        g.min[i] = bounds->min[i];
        g.dim[i] = bounds->dim[i];
        g.sub[i] = sub4[i];
    }
    uint64_t ncells = (uint64_t)g.sub[0] * g.sub[1] * g.sub[2];
    if (g.sub[0] > (uint32_t)kMaxSX)
        return fail(c, ARP_E_LIMIT, "grid has %u cells along x (max %d)", g.sub[0], kMaxSX);
    if (ncells >= 0x7FFFFFFFull) return fail(c, ARP_E_LIMIT, "grid has %llu cells (max 2^31-1)", (unsigned long long)ncells);
    g.nrows = g.sub[1] * g.sub[2];
    g.rowstride = g.sub[0] + 1;
    uint32_t dim = 0;
    for (int i = 0; i < 4; i++)
        if (bounds->dim[i] > 0.0f) dim++;
    g.dim_n = dim;
    c->grid = g;
    c->ncells = ncells;
    c->phys.h = h;
    c->phys.k = arp_norm_const(dim, h);
    c->phys.g = gravity;
    c->phys.eps_hh = 0.1f * h * h;
    c->phys.nm = nm;
    c->n = n;
    c->ns = ns;
    c->nb = nb;
    c->nm = nm;
    c->mats_h.assign(mats, mats + nm);
    c->solids_h.assign(solids, solids + ns);
    c->boundary_h.assign(boundary, boundary + nb);
    c->inter_h.assign(inter, inter + (size_t)nm * nm);
    c->bounds_h = *bounds;
    c->alpha_h = artificial_viscosity;
    c->kscale_h.clear();
    free_buf(c->kscale);

    std::vector<DMat> dm(nm);
    for (uint32_t i = 0; i < nm; i++) {
        const ArpMaterial& m = mats[i];
        DMat d{};
        d.mass = m.mass;
        d.start_den = m.start_den;
        d.c = m.sound_speed;
        d.target_den = m.target_den;
        d.visc = m.visc;
        d.bulk_visc = m.bulk_visc;
        d.lambda = m.normal_stiffness;
        d.mu = m.shear_stiffness;
        d.lambda_d = m.normal_damp;
        d.mu_d = m.shear_damp;
        d.yield = m.yield_strength;
        d.hard = m.work_hardening;
        d.hard_exp = m.work_hardening_exp;
        d.kin_hard = m.kinematic_hardening;
        d.relax = m.relaxation_time;
        d.eq = m.state_eq;
        d.plastic = m.plastic;
        d.elastic = is_stiff(m) ? 1u : 0u;
        d.strain_order = m.strain_order;
        dm[i] = d;
    }
    std::vector<DInter> di((size_t)nm * nm);
    for (size_t i = 0; i < di.size(); i++)
        di[i] = DInter{inter[i].potential, inter[i].strength, inter[i].radius, inter[i].dampening, inter[i].friction};
    int rc;
    if ((rc = upload(c, c->mats, dm.data(), dm.size() * sizeof(DMat)))) return rc;
    if ((rc = upload(c, c->inter, di.data(), di.size() * sizeof(DInter)))) return rc;

    // per-particle static words
    c->mat_by_pid.resize(n);
    c->sid_by_pid.resize(n);
    c->mk_by_pid.resize(n);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t m = particles[i].mat;
        if (m >= nm) return fail(c, ARP_E_ARG, "particle %u has material %u >= %u", i, m, nm);
        uint32_t sid = particles[i].solid_id;
        c->mat_by_pid[i] = m;
        c->sid_by_pid[i] = sid;
        uint32_t mk = m;
        if (sid != ARP_INVALID_INDEX) {
            mk |= kMkSolid;
            if (is_stiff(mats[m])) mk |= kMkElastic;
        }
        c->mk_by_pid[i] = mk;
    }
    if ((rc = upload(c, c->mkStatic, c->mk_by_pid.data(), 4 * (size_t)n))) return rc;
    if ((rc = upload(c, c->matByPid, c->mat_by_pid.data(), 4 * (size_t)n))) return rc;
    c->mobile_mat = n ? (int64_t)c->mat_by_pid[0] : -1;
    c->any_elastic = false;
    for (uint32_t i = 0; i < n; i++)
        if (c->mk_by_pid[i] & kMkElastic) c->any_elastic = true;
    for (uint32_t i = 1; i < n && c->mobile_mat >= 0; i++)
        if (c->mat_by_pid[i] != (uint32_t)c->mobile_mat) c->mobile_mat = -1;
    if ((rc = upload(c, c->sidExtByPid, c->sid_by_pid.data(), 4 * (size_t)n))) return rc;

    // static boundary grid (mode 0 fill, once: neighbor_list.rs:247-264)
    {
        std::vector<uint64_t> key(nb);
        for (uint32_t i = 0; i < nb; i++) {
            uint32_t cx, cy, cz;
            const float p4[4] = {boundary[i].pos[0], boundary[i].pos[1], boundary[i].pos[2], boundary[i].pos[3]};
            bool fin = std::isfinite(p4[3]) && host_cell(g, p4, cx, cy, cz);
            key[i] = fin ? (uint64_t)cx + (uint64_t)g.sub[0] * (cy + (uint64_t)g.sub[1] * cz) : UINT64_MAX;
            if (boundary[i].mat >= nm) return fail(c, ARP_E_ARG, "boundary %u has material %u >= %u", i, boundary[i].mat, nm);
        }
        std::vector<uint32_t> kept;
        cell_sort(ncells, key, {}, c->bP_h, kept);
        c->b_orig = kept;
        std::vector<float> bp(4 * (kept.size() + kPad)), bv(4 * (kept.size() + kPad));
        std::vector<uint32_t> bm(kept.size() + kPad);
        for (size_t s = 0; s < kept.size(); s++) {
            const ArpParticle& b = boundary[kept[s]];
            const ArpMaterial& m = mats[b.mat];
            bp[4 * s + 0] = b.pos[0];
            bp[4 * s + 1] = b.pos[1];
            bp[4 * s + 2] = b.pos[2];
            bp[4 * s + 3] = b.den;
            bv[4 * s + 0] = b.vel[0];
            bv[4 * s + 1] = b.vel[1];
            bv[4 * s + 2] = b.vel[2];
            bv[4 * s + 3] = host_pressure(m, b.den);
            uint32_t mk = b.mat;
            // elastic2 of a boundary neighbour reads particles[boundary id].solid_id
            // (fluid.rs:141, 174): reproduced for the boundary's own material
            uint32_t id = kept[s];
            if (id < n && particles[id].solid_id != ARP_INVALID_INDEX && is_stiff(m)) mk |= kMkElastic;
            bm[s] = mk;
        }
        c->bnd_mat = -1;
        c->bnd_elastic = false;
        for (size_t s = 0; s < kept.size(); s++) {
            const int64_t m = (int64_t)(bm[s] & kMatMask);
            if (bm[s] & kMkElastic) c->bnd_elastic = true;
            if (c->bnd_mat == -1) c->bnd_mat = m;
            else if (c->bnd_mat != m) c->bnd_mat = -2;
        }
        if ((rc = upload(c, c->bposd, bp.data(), bp.size() * 4))) return rc;
        if ((rc = upload(c, c->bvelq, bv.data(), bv.size() * 4))) return rc;
        if ((rc = upload(c, c->bmk, bm.data(), bm.size() * 4))) return rc;
        if ((rc = upload(c, c->bP, c->bP_h.data(), c->bP_h.size() * 4))) return rc;
    }
    // static reference grid (mode 1 fill: neighbor_list.rs:266-275), capped at 128 - count[0]
    {
        std::vector<uint64_t> key(n, UINT64_MAX);
        for (uint32_t i = 0; i < n; i++) {
            uint32_t sid = particles[i].solid_id;
            if (sid == ARP_INVALID_INDEX || sid >= ns) continue;
            uint32_t cx, cy, cz;
            const float* p4 = solids[sid].ref_pos;
            if (std::isfinite(p4[3]) && host_cell(g, p4, cx, cy, cz))
                key[i] = (uint64_t)cx + (uint64_t)g.sub[0] * (cy + (uint64_t)g.sub[1] * cz);
        }
        std::vector<uint32_t> room(ncells);
        for (uint64_t cc = 0; cc < ncells; cc++) room[cc] = kBucketCap - (c->bP_h[cc + 1] - c->bP_h[cc]);
        std::vector<uint32_t> kept;
        cell_sort(ncells, key, room, c->rP_h, kept);
        if ((rc = upload(c, c->rP, c->rP_h.data(), c->rP_h.size() * 4))) return rc;
        // the internal solid order: the solids of the reference grid in its cell order, then
        // the others (refused by a full cell, non-finite, unused) by id
        c->s_ord.clear();
        c->s_rank.assign(ns, ARP_INVALID_INDEX);
        for (uint32_t p : kept) {
            const uint32_t sid = particles[p].solid_id;
            if (sid < ns && c->s_rank[sid] == ARP_INVALID_INDEX) {
                c->s_rank[sid] = (uint32_t)c->s_ord.size();
                c->s_ord.push_back(sid);
            }
        }
        for (uint32_t sid = 0; sid < ns; sid++)
            if (c->s_rank[sid] == ARP_INVALID_INDEX) {
                c->s_rank[sid] = (uint32_t)c->s_ord.size();
                c->s_ord.push_back(sid);
            }
        {
            std::vector<uint32_t> si(std::max<uint32_t>(n, 1), ARP_INVALID_INDEX);
            for (uint32_t i = 0; i < n; i++) {
                const uint32_t sid = c->sid_by_pid[i];
                si[i] = sid < ns ? c->s_rank[sid] : sid;
            }
            if ((rc = upload(c, c->sidByPid, si.data(), 4 * (size_t)n))) return rc;
            if ((rc = upload(c, c->sOrd, c->s_ord.data(), 4 * (size_t)std::max<uint32_t>(ns, 1)))) return rc;
        }
        // the solid path runs only when there is more than one solid (shaders/mod.rs:51)
        c->solids_active = ns > 1;
        if (c->solids_active)
            if ((rc = build_solid_lists(c, particles, n, solids, ns, mats, kept))) return rc;
    }
    if ((rc = alloc_state(c))) return rc;
    c->nl_thresh = -1;
    c->nl_dirty = true;
    c->have_scene = true;
    c->cur = 0;
    c->time = 0.0;
    c->keys_valid = false;
    c->have_build = false;
    if ((rc = fill_u32(c, c->binCount, g.nrows + 1, 0))) return rc;
    if ((rc = arp_upload_state(c, particles, n, solids, ns))) return rc;
    return arp_synchronize(c);
}

int arp_upload_state(arp_ctx* c, const ArpParticle* particles, uint32_t n, const ArpSolidParticle* solids, uint32_t ns) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (n != c->n || (n && !particles)) return fail(c, ARP_E_ARG, "state has %u particles, scene %u", n, c->n);
    c->rec_valid = false;
    c->nl_dirty = true;
    c->rec_prefilled = nullptr;
    c->verlet_valid = false;
    c->yland_stale = false;
    HIPCHK(c, hipMemcpyAsync(c->aos_tmp.p, particles, sizeof(ArpParticle) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    if (c->solids_active) {
        const ArpSolidParticle* src = solids ? solids : c->solids_h.data();
        if (solids && ns != c->ns) return fail(c, ARP_E_ARG, "state has %u solids, scene %u", ns, c->ns);
        if ((rc = upload(c, c->saos_tmp, src, sizeof(ArpSolidParticle) * (size_t)c->ns))) return rc;
        EpsInArgs e{c->ns, c->saos_tmp.p, P<float>(c->Yeps), P<uint32_t>(c->sOrd)};
        if ((rc = launch(c, K_EPS_IN, nblk(c->ns), 256, e))) return rc;
    }
    if (c->keys_valid)
        if ((rc = fill_u32(c, c->binCount, c->grid.nrows + 1, 0))) return rc;
    if ((rc = aos_to_soa(c, c->aos_tmp.p, P<float>(c->Yland[0]), P<float>(c->Yland[1]),
                         P<uint32_t>(c->meta_mk[c->cur]), P<uint32_t>(c->meta_pid[c->cur]))))
        return rc;
    c->keys_valid = false;
    return ARP_OK;
}

int arp_set_integrator(arp_ctx* c, const ArpTableau* t, float timestep, uint32_t subticks) {
    if (!c || !t) return ARP_E_ARG;
    c->verlet_valid = false;
    if (t->stages == ARP_VERLET) {
        c->tab = *t;
        c->timestep = timestep;
        c->subticks = subticks ? subticks : 1;
        return ARP_OK;
    }
    if (t->stages < 1 || t->stages > ARP_MAX_STAGES) return fail(c, ARP_E_ARG, "tableau has %d stages", t->stages);
    for (int i = 0; i < t->stages; i++)
        for (int j = i; j < t->stages; j++)
            if (t->a[i][j] != 0.0f) return fail(c, ARP_E_ARG, "tableau is not explicit (a[%d][%d] != 0)", i, j);
    c->tab = *t;
    c->timestep = timestep;
    c->subticks = subticks ? subticks : 1;
    return ARP_OK;
}

// Steps of a plain list-keeping multi-stage method are issued speculatively: each stage as
// its walk kernel alone (no rebuild chain), kSpecW steps in flight.  A stage whose lists are
// not valid records the step in a host-mapped word and every later stage returns at once;
// the host then restores the list words the step began with and runs that step with the
// rebuild chain.  The result is the one of issuing the chain at every stage.  The call
// returns once its steps have run, so calls of fewer than kSpecMin steps (one frame at a
// time, say) stay asynchronous and issue the chain.  The host waits on an event after every
// kSpecE steps (an event between two kernels idles the GPU for about 6 us, so groups of
// eight steps: one such pause per 16 launches of a two-stage method); a snapshot slot is
// overwritten only by a later step, which returns at once when an earlier one aborted, so
// the slot of an aborted step survives until its restore.
constexpr uint32_t kSpecW = 16;
constexpr uint32_t kSpecE = 8;
constexpr uint32_t kSpecMin = 8;
static_assert(kSpecW % kSpecE == 0 && kSpecW / kSpecE < 4, "events in flight: at most three of four");

static bool spec_ok(arp_ctx* c) {
    const ArpTableau& T = c->tab;
    return !c->nl_nospec && !c->capture && !c->nl_always && c->nl_max_age == 0 && T.stages > 1 &&
           nl_enabled(c) && !c->nl_dirty && c->nlWords.p;
}

// The host-mapped abort word and the events; false (and speculation off for the context)
// when the runtime cannot provide them.
static bool spec_setup(arp_ctx* c) {
    if (!c->hAbort) {
        bool ok = hipHostMalloc((void**)&c->hAbort, 64, hipHostMallocMapped) == hipSuccess;
        ok = ok && hipHostGetDevicePointer((void**)&c->hAbortDev, c->hAbort, 0) == hipSuccess;
        for (hipEvent_t& e : c->specEv) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            c->nl_nospec = true;
            return false;
        }
    }
    if (dalloc(c, c->nlSnap, 4 * (size_t)kSpecSnap * kSpecW) != ARP_OK) {
        c->nl_nospec = true;
        return false;
    }
    return true;
}

// Steps [j0, j0 + count) of a recording of nsteps steps, into the recorded steps j % K;
// advance: whether the steps advance the clock (a replay of a segment does not).
struct RecRange {
    uint32_t j0, nsteps, K;
    bool advance;
};

// Step k of a run (of a recording, when rr is given).
static int issue_step(arp_ctx* c, float dt, const RecRange* rr, uint32_t k, const SpecArg* sp) {
    if (!rr) return step_once(c, dt, nullptr, sp);
    const uint32_t j = rr->j0 + k;
    c->rec_next = j + 1 < rr->nsteps ? &c->rec[(j + 1) % rr->K] : nullptr;
    const int rc = step_once(c, dt, &c->rec[j % rr->K], sp);
    c->rec_next = nullptr;
    return rc;
}

// Runs nsteps steps; a run of kSpecMin steps or more on kept lists is issued speculatively,
// kSpecW steps in flight, and a step whose lists turn out invalid runs again with the chain.
// A recorded step speculates when the previous one left its input in its stage buffers.
// (spec_min: the fewest steps issued speculatively; the reverse-mode calls, which wait for
// their steps anyway, speculate on runs of two)
static int run_steps(arp_ctx* c, float dt, uint32_t nsteps, const RecRange* rr = nullptr,
                     uint32_t spec_min = kSpecMin) {
    int rc;
    uint32_t k = 0;
    const bool adv = !rr || rr->advance;
    while (k < nsteps) {
        if (rr && rr->K >= 2 && nsteps - k >= spec_min && spec_ok(c)) {
            // a recorded step speculates from its stage buffers: bring its input there
            // (what the chain's reusing sort copies; a rebuild sorts from the landing ones)
            RecStep& R = c->rec[(rr->j0 + k) % rr->K];
            if (c->rec_prefilled != R.Xp[0].p) {
                if ((rc = sync_yland(c))) return rc;
                const size_t bytes = 16 * (size_t)c->n;
                HIPCHK(c, hipMemcpyAsync(R.Xp[0].p, c->Yland[0].p, bytes, hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(c, hipMemcpyAsync(R.Xv[0].p, c->Yland[1].p, bytes, hipMemcpyDeviceToDevice, c->stream));
                c->rec_prefilled = R.Xp[0].p;
            }
        }
        if (nsteps - k < spec_min || !spec_ok(c) ||
            (rr && (rr->K < 2 || c->rec_prefilled != c->rec[(rr->j0 + k) % rr->K].Xp[0].p))) {
            if ((rc = issue_step(c, dt, rr, k, nullptr))) return rc;
            if (adv) c->time += dt;
            k++;
            continue;
        }
        if (!spec_setup(c)) continue;  // (the steps are then issued with the rebuild chain)
        __atomic_store_n(c->hAbort, kInvalid, __ATOMIC_RELEASE);
        // the list counter the speculative stages continue from
        uint32_t gen0 = 0;
        HIPCHK(c, hipMemcpyAsync(&gen0, P<uint32_t>(c->nlWords) + 10, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const uint32_t S = (uint32_t)c->tab.stages;
        struct Host {
            int cur;
            bool keys_valid, have_build, yland_stale;
            const void* prefilled;
        } hs[kSpecW];
        const uint32_t m = nsteps - k;
        uint32_t issued = 0, done = 0;
        bool aborted = false;
        while (done < m) {
            while (issued < m && issued - done < kSpecW) {
                hs[issued % kSpecW] = Host{c->cur, c->keys_valid, c->have_build, c->yland_stale, c->rec_prefilled};
                const SpecArg sa{issued, issued % kSpecW, gen0 + issued * S};
                if ((rc = issue_step(c, dt, rr, k + issued, &sa))) {
                    hipStreamSynchronize(c->stream);
                    return rc;
                }
                issued++;
                if (issued % kSpecE == 0 || issued == m)  // the end of a group of kSpecE steps
                    HIPCHK(c, hipEventRecord(c->specEv[((issued - 1) / kSpecE) % 4], c->stream));
            }
            // (done is at a group's start, and issued at least kSpecW further or at m, so the
            // group's event is recorded)
            const uint32_t grp = done / kSpecE, grpEnd = std::min(m, (grp + 1) * kSpecE);
            HIPCHK(c, hipEventSynchronize(c->specEv[grp % 4]));
            if (__atomic_load_n(c->hAbort, __ATOMIC_ACQUIRE) != kInvalid) {
                aborted = true;
                break;
            }
            for (; done < grpEnd; done++) {
                if (adv) c->time += dt;
                c->spec_steps++;
            }
        }
        if (!aborted) {
            k += m;
            continue;
        }
        // every issued step has run (those after the abort did nothing)
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const uint32_t code = __atomic_load_n(c->hAbort, __ATOMIC_ACQUIRE);
        const uint32_t a = code >> 4, stage = code & 15u;
        if (a < done || a >= issued) return fail(c, ARP_E_STATE, "speculative step %u out of range", a);
        for (; done < a; done++) {
            if (adv) c->time += dt;
            c->spec_steps++;
        }
        const Host& h = hs[a % kSpecW];
        c->cur = h.cur;
        c->keys_valid = h.keys_valid;
        c->have_build = h.have_build;
        c->yland_stale = h.yland_stale;
        c->rec_prefilled = h.prefilled;
        // the list words as the step found them (a first stage that aborted changed nothing);
        // a recorded step's stages advanced the build stamps of its own slots
        SpecRestoreArgs sr{P<uint32_t>(c->nlWords),
                           stage > 0 ? P<uint32_t>(c->nlSnap) + (size_t)kSpecSnap * (a % kSpecW) : nullptr,
                           c->slot.stamp, P<uint32_t>(c->stats), nullptr, 0u};
        if (rr) {
            const RecStep& R = c->rec[(rr->j0 + k + a) % rr->K];
            sr.stamp = R.slot[0].stamp;
            if (stage > 1) {
                sr.stampDec = R.slot[1].stamp;  // (the slots' stamps are consecutive words)
                sr.ndec = stage - 1;
            }
        }
        if ((rc = launch(c, K_SPEC_RESTORE, 1, 64, sr))) return rc;
        __atomic_store_n(c->hAbort, kInvalid, __ATOMIC_RELEASE);
        c->spec_aborts++;
        k += a;
        // that step with the rebuild chain
        if ((rc = issue_step(c, dt, rr, k, nullptr))) return rc;
        if (adv) c->time += dt;
        k++;
    }
    return ARP_OK;
}

int arp_step(arp_ctx* c, float dt, uint32_t nsteps) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    return run_steps(c, dt, nsteps);
}

int arp_update(arp_ctx* c, uint32_t nframes) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    const float dt = c->timestep / (float)c->subticks;
    HIPCHK(c, hipSetDevice(c->device));
    return run_steps(c, dt, nframes * c->subticks);
}

double arp_time(const arp_ctx* c) { return c ? c->time : 0.0; }

int arp_adjoint_record(arp_ctx* c, float dt, uint32_t nsteps, uint32_t checkpoint_every) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (c->tab.stages == ARP_VERLET)
        return fail(c, ARP_E_STATE, "reverse mode covers the Runge-Kutta tableaus (not velocity Verlet)");
    if (nsteps == 0 || checkpoint_every == 0 || !(dt > 0.0f) || !std::isfinite(dt))
        return fail(c, ARP_E_ARG, "adjoint_record needs nsteps >= 1, checkpoint_every >= 1 and dt > 0");
    if (!nl_fits_u32(rec_cap(c), c->n))
        return fail(c, ARP_E_LIMIT, "the neighbour record of %u particles exceeds 4 GB per stage", c->n);
    // every step of this recording, and every replay of it, keeps this many neighbours
    c->rec_nb_cap = rec_cap(c);
    HIPCHK(c, hipSetDevice(c->device));
    const uint32_t K = std::min(checkpoint_every, nsteps);
    const uint32_t nseg = (nsteps + K - 1) / K;
    const int S = c->tab.stages;
    trim_rec(c, K, nseg + 1);
    c->rec.resize(K);
    c->ck.resize(nseg + 1);  // segment starts, then the end state
    for (RecStep& R : c->rec)
        if ((rc = alloc_rec(c, R, S))) return rc;
    // every step's stage flags in one buffer (read back with one copy)
    if ((rc = dalloc(c, c->recFlags, 4 * 2 * ARP_MAX_STAGES * (size_t)K))) return rc;
    for (uint32_t j = 0; j < K; j++) c->rec[j].fl = P<uint32_t>(c->recFlags) + 2 * ARP_MAX_STAGES * (size_t)j;
    for (uint32_t seg = 0; seg < nseg; seg++) {
        const uint32_t j0 = seg * K, j1 = std::min(nsteps, j0 + K);
        if ((rc = ck_save(c, c->ck[seg], nseg > 1))) return rc;
        // kept neighbour lists restart with the recorded segment when the reverse sweep
        // replays segments (a second sweep replays it too), as every replay from a checkpoint
        // does, so that every replay repeats the recorded steps exactly.  The segments run
        // only for their checkpoints keep their lists (a replay then differs from them only in
        // the order of the per-particle sums)
        if (nseg > 1 && seg + 1 == nseg) c->nl_dirty = true;
        // only the last segment is kept; the others run as plain steps (the same arithmetic)
        // for their checkpoints, and the reverse sweep replays them
        const RecRange rr{j0, nsteps, K, true};
        if ((rc = run_steps(c, dt, j1 - j0, seg + 1 == nseg ? &rr : nullptr, 2))) return rc;
    }
    c->rec_prefilled = nullptr;
    if ((rc = ck_save(c, c->ck[nseg], nseg > 1))) return rc;
    if ((rc = fetch_over(c, (nsteps - 1) % K + 1))) return rc;
    c->rec_dt = dt;
    c->rec_nsteps = nsteps;
    c->rec_K = K;
    c->rec_seg = nseg - 1;
    c->rec_tab = c->tab;
    c->rec_valid = true;
    return ARP_OK;
}

int arp_set_stiffness_field(arp_ctx* c, const float* scale, uint32_t ns) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (!scale) {
        free_buf(c->kscale);
        c->kscale_h.clear();
        return ARP_OK;
    }
    if (ns != c->ns) return fail(c, ARP_E_ARG, "stiffness field has %u entries, scene %u solids", ns, c->ns);
    c->kscale_h.assign(scale, scale + ns);
    c->rec_valid = false;
    std::vector<float> ks(std::max<uint32_t>(ns, 1), 1.0f);
    for (uint32_t t = 0; t < ns; t++) ks[t] = scale[c->s_ord[t]];
    return upload(c, c->kscale, ks.data(), 4 * ks.size());
}

int arp_adjoint_stiffness_grad(arp_ctx* c, float* out, uint32_t ns) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (!c->solids_active || !c->spar.p) return fail(c, ARP_E_STATE, "no solid reverse sweep has run");
    if (!out || ns != c->ns) return fail(c, ARP_E_ARG, "need %u entries", c->ns);
    std::vector<float> sp(4 * (size_t)ns);
    HIPCHK(c, hipMemcpyAsync(sp.data(), c->spar.p, 4 * sp.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // lambda_eff = k_s lambda, mu_eff = k_s mu: dL/dk_s = lambda dL/dlambda_eff + mu dL/dmu_eff
    for (uint32_t s = 0; s < ns; s++) {
        const uint32_t p = c->solids_h[s].part_id;
        if (p >= c->n) {
            out[s] = 0.0f;
            continue;
        }
        const ArpMaterial& M = c->mats_h[c->mat_by_pid[p]];
        const size_t t = c->s_rank[s];
        out[s] = M.normal_stiffness * sp[4 * t] + M.shear_stiffness * sp[4 * t + 1];
    }
    return ARP_OK;
}

int arp_adjoint_backward(arp_ctx* c, const void* d_grad_out, void* d_grad_in, float* grad_params) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (!c->rec_valid) return fail(c, ARP_E_STATE, "no recorded trajectory (call arp_adjoint_record)");
    if (!d_grad_out || !d_grad_in) return fail(c, ARP_E_ARG, "null cotangent buffer");
    if (grad_params && c->nm > (uint32_t)kMatLds)
        return fail(c, ARP_E_LIMIT, "material parameter gradients need at most %d materials", kMatLds);
    if (memcmp(&c->rec_tab, &c->tab, sizeof(ArpTableau)) != 0)
        return fail(c, ARP_E_STATE, "the integrator changed after the trajectory was recorded");
    if (rec_cap(c) != c->rec_nb_cap)  // replays write the recorded lists
        return fail(c, ARP_E_STATE, "the neighbour settings changed after the trajectory was recorded");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t n = std::max<uint32_t>(c->n, 1) + kPad;
    const int S = c->tab.stages;
    for (int j = 0; j < 2; j++) {
        if ((rc = dalloc(c, c->LamIn[j], 16 * n)) || (rc = dalloc(c, c->LamAcc[j], 16 * n)) ||
            (rc = dalloc(c, c->LamNext[j], 16 * n)) || (rc = dalloc(c, c->Wb[j], 16 * n)))
            return rc;
        for (int s = 0; s < S; s++)
            if ((rc = dalloc(c, c->Gb[s][j], 16 * n))) return rc;
    }
    if ((rc = dalloc(c, c->pgradPart, 4 * (size_t)nblk(c->n) * 2 * kMatLds))) return rc;
    if ((rc = fill_u32(c, c->pgradPart, nblk(c->n) * 2 * kMatLds, 0))) return rc;  // accumulated per block
    if (c->solids_active) {
        const size_t ns = std::max<uint32_t>(c->ns, 1);
        // (20 floats per solid: the kernels' padded rows of 18 F, F' cotangents)
        if ((rc = dalloc(c, c->Fbar, 4 * 20 * ns)) || (rc = dalloc(c, c->spar, 4 * 4 * ns))) return rc;
        if (any_other_order(c) && (rc = dalloc(c, c->Ebuf, 4 * 48 * ns))) return rc;
        if ((rc = fill_u32(c, c->spar, 4 * c->ns, 0))) return rc;
        if (any_plastic(c)) {  // the plastic strain's cotangents: dL/d y_N = 0 (the loss is on particles)
            if ((rc = dalloc(c, c->LamE, 64 * ns))) return rc;
            for (int s = 0; s < S; s++)
                if ((rc = dalloc(c, c->GbE[s], 64 * ns))) return rc;
            if ((rc = fill_u32(c, c->LamE, 16 * c->ns, 0))) return rc;
        }
    }
    if ((rc = dalloc(c, c->pgradSum, 4 * 2 * kMatLds))) return rc;
    if ((rc = fill_u32(c, c->pgradSum, 2 * kMatLds, 0))) return rc;
    const uint32_t K = c->rec_K, nsteps = c->rec_nsteps, nseg = (nsteps + K - 1) / K;
    Checkpoint& fin = c->ck[nseg];
    // dL/dy_N (records by particle id) -> the final storage order
    GradAosArgs in{c->n, 0, P<uint32_t>(fin.pid), const_cast<void*>(d_grad_out), P<float>(c->LamIn[0]),
                   P<float>(c->LamIn[1])};
    if ((rc = launch(c, K_GRAD_AOS, nblk(c->n), 256, in))) return rc;
    bool replayed = false;
    for (int seg = (int)nseg - 1; seg >= 0; seg--) {
        const uint32_t j0 = (uint32_t)seg * K, j1 = std::min(nsteps, j0 + K);
        if ((uint32_t)seg != c->rec_seg) {
            // replay the segment from its checkpoint, recording
            replayed = true;
            if ((rc = ck_restore(c, c->ck[seg]))) return rc;
            const RecRange rr{j0, j1, K, false};
            if ((rc = run_steps(c, c->rec_dt, j1 - j0, &rr, 2))) return rc;
            c->rec_prefilled = nullptr;
            if ((rc = fetch_over(c, j1 - j0))) return rc;
            c->rec_seg = (uint32_t)seg;
        }
        for (uint32_t j = j1; j-- > j0;)
            if ((rc = reverse_step(c, c->rec[j - j0], c->rec_dt))) return rc;
    }
    if (c->nm <= (uint32_t)kMatLds) {  // the per-block parameter partials of the whole sweep
        ReduceArgs r{nblk(c->n), 2 * c->nm, P<float>(c->pgradPart), P<float>(c->pgradSum)};
        if ((rc = launch(c, K_REDUCE, 2 * c->nm, 256, r))) return rc;
    }
    // dL/dy_0 is now in the storage order of the recorded start state
    GradAosArgs out{c->n, 1, P<uint32_t>(c->ck[0].pid), d_grad_in, P<float>(c->LamIn[0]), P<float>(c->LamIn[1])};
    if ((rc = launch(c, K_GRAD_AOS, nblk(c->n), 256, out))) return rc;
    // the live state goes back to the end of the recorded trajectory (a sweep without replays
    // left it there, with its kept neighbour lists)
    if (replayed && (rc = ck_restore(c, fin))) return rc;
    if (grad_params) {
        std::vector<float> pg(2 * (size_t)c->nm);
        HIPCHK(c, hipMemcpyAsync(pg.data(), c->pgradSum.p, 4 * pg.size(), hipMemcpyDeviceToHost, c->stream));
        std::vector<float> sp;
        if (c->solids_active) {
            sp.resize(4 * (size_t)c->ns);
            HIPCHK(c, hipMemcpyAsync(sp.data(), c->spar.p, 4 * sp.size(), hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        std::vector<double> acc(6 * (size_t)c->nm, 0.0);
        for (uint32_t m = 0; m < c->nm; m++) {
            acc[6 * m + 0] = pg[2 * m];
            acc[6 * m + 1] = pg[2 * m + 1];
        }
        // per-solid partials w.r.t. the effective (scaled) constants -> per material
        for (uint32_t t = 0; t < c->ns && !sp.empty(); t++) {  // (internal order)
            const uint32_t s = c->s_ord[t];
            const uint32_t p = c->solids_h[s].part_id;
            if (p >= c->n) continue;
            const uint32_t m = c->mat_by_pid[p];
            const float ks = c->kscale_h.empty() ? 1.0f : c->kscale_h[s];
            acc[6 * m + 2] += (double)sp[4 * t] * ks;
            acc[6 * m + 3] += (double)sp[4 * t + 1] * ks;
            acc[6 * m + 4] += sp[4 * t + 2];
            acc[6 * m + 5] += sp[4 * t + 3];
        }
        for (size_t q = 0; q < acc.size(); q++) grad_params[q] = (float)acc[q];
    }
    return ARP_OK;
}

int arp_download(arp_ctx* c, ArpParticle* particles, uint32_t n, ArpSolidParticle* solids, uint32_t ns) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (particles) {
        if (n != c->n) return fail(c, ARP_E_ARG, "download of %u particles, scene has %u", n, c->n);
        if ((rc = sync_yland(c))) return rc;
        if ((rc = soa_to_aos(c, P<float>(c->Yland[0]), P<float>(c->Yland[1]), P<uint32_t>(c->meta_pid[c->cur]), c->aos_tmp.p)))
            return rc;
        HIPCHK(c, hipMemcpyAsync(particles, c->aos_tmp.p, sizeof(ArpParticle) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    }
    if (solids) {
        if (ns != c->ns) return fail(c, ARP_E_ARG, "download of %u solids, scene has %u", ns, c->ns);
        if (c->solids_active) {
            if ((rc = dalloc(c, c->saos_tmp, sizeof(ArpSolidParticle) * (size_t)ns))) return rc;
            EpsOutArgs e{ns, P<float>(c->Yeps), P<float>(c->ref_d), P<uint32_t>(c->part_d), c->saos_tmp.p,
                         P<uint32_t>(c->sOrd)};
            if ((rc = launch(c, K_EPS_OUT, nblk(ns), 256, e))) return rc;
            HIPCHK(c, hipMemcpyAsync(solids, c->saos_tmp.p, sizeof(ArpSolidParticle) * (size_t)ns,
                                     hipMemcpyDeviceToHost, c->stream));
        } else {
            std::copy(c->solids_h.begin(), c->solids_h.end(), solids);
        }
    }
    return arp_synchronize(c);
}

namespace {

// Material equality of the reference's derived PartialEq: field by field, floats by value.
bool same_material(const ArpMaterial& a, const ArpMaterial& b) {
    return a.immobile == b.immobile && a.mass == b.mass && a.start_den == b.start_den && a.state_eq == b.state_eq &&
           a.sound_speed == b.sound_speed && a.target_den == b.target_den && a.visc == b.visc &&
           a.bulk_visc == b.bulk_visc && a.strain_order == b.strain_order &&
           a.normal_stiffness == b.normal_stiffness && a.shear_stiffness == b.shear_stiffness &&
           a.normal_damp == b.normal_damp && a.shear_damp == b.shear_damp && a.plastic == b.plastic &&
           a.yield_strength == b.yield_strength && a.work_hardening == b.work_hardening &&
           a.work_hardening_exp == b.work_hardening_exp && a.kinematic_hardening == b.kinematic_hardening &&
           a.thermal_softening == b.thermal_softening && a.relaxation_time == b.relaxation_time;
}

}  // namespace

float arp_kernel_radius(const arp_ctx* c) { return c && c->have_scene ? c->phys.h : 0.0f; }

uint32_t arp_num_materials(const arp_ctx* c) { return c && c->have_scene ? c->nm : 0u; }

uint32_t arp_num_particles(const arp_ctx* c) { return c && c->have_scene ? c->n : 0u; }

uint32_t arp_num_solids(const arp_ctx* c) { return c && c->have_scene ? c->ns : 0u; }

int arp_add_particles(arp_ctx* c, const ArpMaterial* material, const ArpParticle* particles, uint32_t count,
                      uint32_t* material_added) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (!material || (count && !particles)) return fail(c, ARP_E_ARG, "arp_add_particles: missing argument");
    if (material_added) *material_added = 0;
    if (count == 0) return ARP_OK;  // Particles::add_particles returns early
    if ((uint64_t)c->n + count > 0xFFFFFFF0ull) return fail(c, ARP_E_LIMIT, "more than 2^32 particles");
    const uint32_t n = c->n, nm = c->nm;
    // the live state in the caller's order (plastic strain included)
    std::vector<ArpParticle> P(n + (size_t)count);
    std::vector<ArpSolidParticle> S(c->ns);
    if ((rc = arp_download(c, P.data(), n, S.data(), c->ns))) return rc;
    // the material: a stiff one is always new, else the first equal one (particles.rs:549-565)
    const bool solid = is_stiff(*material);
    uint32_t mid = nm;
    if (!solid)
        for (uint32_t i = 0; i < nm; i++)
            if (same_material(c->mats_h[i], *material)) {
                mid = i;
                break;
            }
    std::vector<ArpMaterial> M = c->mats_h;
    std::vector<ArpMatInteraction> I = c->inter_h;
    const float h = c->phys.h;
    if (mid == nm) {
        // the grown matrix (particles.rs:573-598): old pairs kept, default_between for the
        // new row and column, and for the new material with itself unless it is stiff
        if (nm + 1 > 0xFFFF) return fail(c, ARP_E_LIMIT, "at most 65535 materials");
        M.push_back(*material);
        const uint32_t m2 = nm + 1;
        std::vector<ArpMatInteraction> J((size_t)m2 * m2);
        for (uint32_t i = 0; i < nm; i++) {
            for (uint32_t j = 0; j < nm; j++) J[(size_t)i * m2 + j] = I[(size_t)i * nm + j];
            arp_default_interaction(&M[i], &M[nm], h, &J[(size_t)i * m2 + nm]);
        }
        for (uint32_t i = 0; i < nm; i++) arp_default_interaction(&M[nm], &M[i], h, &J[(size_t)nm * m2 + i]);
        if (solid)
            J[(size_t)nm * m2 + nm] = ArpMatInteraction{0, 0.0f, 0.0f, 0.0f, 0.0f};
        else
            arp_default_interaction(&M[nm], &M[nm], h, &J[(size_t)nm * m2 + nm]);
        I.swap(J);
        if (material_added) *material_added = 1;
    }
    // the new particles (particles.rs:601-619); a scene without solids holds one placeholder
    // solid, which new solids replace
    if (solid && S.size() == 1 && S[0].part_id == ARP_INVALID_INDEX) S.clear();
    for (uint32_t k = 0; k < count; k++) {
        ArpParticle q = particles[k];
        q.mat = mid;
        q.solid_id = ARP_INVALID_INDEX;
        if (solid) {
            q.solid_id = (uint32_t)S.size();
            ArpSolidParticle sp{};
            sp.part_id = n + k;
            std::memcpy(sp.ref_pos, q.pos, sizeof sp.ref_pos);
            S.push_back(sp);
        }
        P[n + k] = q;
    }
    if (S.empty()) {
        ArpSolidParticle sp{};
        sp.part_id = ARP_INVALID_INDEX;
        S.push_back(sp);
    }
    std::vector<float> ks = c->kscale_h;
    if (!ks.empty()) ks.resize(S.size(), 1.0f);
    const std::vector<ArpParticle> B = c->boundary_h;
    const ArpAABB bounds = c->bounds_h;
    const double t = c->time;
    if ((rc = arp_scene_upload(c, P.data(), (uint32_t)P.size(), S.data(), (uint32_t)S.size(), B.data(),
                               (uint32_t)B.size(), M.data(), (uint32_t)M.size(), I.data(), &bounds, h, c->phys.g,
                               c->alpha_h)))
        return rc;
    c->time = t;
    if (!ks.empty() && (rc = arp_set_stiffness_field(c, ks.data(), (uint32_t)ks.size()))) return rc;
    return ARP_OK;
}

int arp_neighbors_update(arp_ctx* c, const ArpParticle* d_state) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (!d_state) return fail(c, ARP_E_ARG, "null state");
    return stage_sort_aos(c, d_state);
}

int arp_compute_forces(arp_ctx* c, const ArpParticle* d_state, const ArpSolidParticle* d_solids,
                       ArpParticle* d_deriv, ArpSolidParticle* d_solid_deriv) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (!d_state || !d_deriv) return fail(c, ARP_E_ARG, "null state/derivative");
    if (c->solids_active && !d_solids) return fail(c, ARP_E_ARG, "the scene has solids: d_solids is required");
    if ((rc = stage_sort_aos(c, d_state))) return rc;
    DBuf* kd = c->solids_active ? c->KF : c->K[0];
    FluidArgs f = fluid_base(c);
    f.Xposd = P<float>(c->Xb[1][0]);
    f.Xvelq = P<float>(c->Xb[1][1]);
    f.mk = P<uint32_t>(c->ps_mk[1]);
    f.mode = 2;
    f.KdP = P<float>(kd[0]);
    f.KdV = P<float>(kd[1]);
    if ((rc = launch(c, K_FLUID, nblk(c->n), 256, f))) return rc;
    if (c->solids_active) {
        EpsInArgs e{c->ns, d_solids, P<float>(c->ps_eps), P<uint32_t>(c->sOrd)};
        if ((rc = launch(c, K_EPS_IN, nblk(c->ns), 256, e))) return rc;
        SolidArgs sa = solid_base(c);
        sa.Xposd = P<float>(c->Xb[1][0]);
        sa.Xvelq = P<float>(c->Xb[1][1]);
        sa.mk = P<uint32_t>(c->ps_mk[1]);
        sa.pid = P<uint32_t>(c->ps_pid[1]);
        sa.Xeps = P<float>(c->ps_eps);
        sa.mode = 2;
        sa.KdP = P<float>(kd[0]);
        sa.KdV = P<float>(kd[1]);
        if ((rc = launch(c, K_STRAIN, nblk(c->ns), 256, sa))) return rc;
        if ((rc = launch(c, K_SOLID, nblk(c->ns), 256, sa))) return rc;
    }
    if ((rc = soa_to_aos(c, P<float>(kd[0]), P<float>(kd[1]), P<uint32_t>(c->ps_pid[1]), d_deriv))) return rc;
    if (d_solid_deriv && d_solids) {
        if (c->solids_active) {
            // dest.solids (cleared) + dest2.solids {part_id, 0, plastic rate}
            EpsOutArgs e{c->ns, P<float>(c->keps), nullptr, P<uint32_t>(c->part_d), d_solid_deriv,
                         P<uint32_t>(c->sOrd)};
            if ((rc = launch(c, K_EPS_OUT, nblk(c->ns), 256, e))) return rc;
        } else {
            ClearSolidsArgs a{c->ns, d_solids, d_solid_deriv};
            if ((rc = launch(c, K_CLEAR_SOLIDS, nblk(c->ns), 256, a))) return rc;
        }
    }
    return ARP_OK;
}

// One launch of lin_comb_N / s_lin_comb_N over at most 5 terms; solids take part only when
// both their terms and their output are given.
static int lincomb5(arp_ctx* c, uint32_t nterms, const float* coefs, const ArpParticle* const* d_terms,
                    const ArpSolidParticle* const* d_solid_terms, ArpParticle* d_out, ArpSolidParticle* d_solid_out) {
    const bool sol = d_solid_terms && d_solid_out;
    LincombArgs a{};
    a.n = c->n;
    a.ns = sol ? c->ns : 0;
    a.nterms = nterms;
    for (uint32_t k = 0; k < nterms; k++) {
        a.coef[k] = coefs[k];
        a.terms[k] = d_terms[k];
        a.sterms[k] = sol ? d_solid_terms[k] : nullptr;
    }
    a.out = d_out;
    a.sout = sol ? d_solid_out : nullptr;
    return launch(c, K_LINCOMB, nblk(std::max(c->n, a.ns)), 256, a);
}

// reduce_owned/reduce_ref (particles.rs:381-417): terms in chunks of 5, each chunk combined
// into a partial, and the partials combined again with coefficient 1.  part[k] / spart[k]
// receive chunk k (spart is used only when solids take part).
static int lincomb_chunked(arp_ctx* c, uint32_t nterms, const float* coefs, const ArpParticle* const* d_terms,
                           const ArpSolidParticle* const* d_solid_terms, ArpParticle* d_out,
                           ArpSolidParticle* d_solid_out, ArpParticle* const* part, ArpSolidParticle* const* spart) {
    const bool sol = d_solid_terms && d_solid_out;
    const uint32_t nch = (nterms + 4) / 5;
    const ArpParticle* tp[5];
    const ArpSolidParticle* sp[5];
    const float ones[5] = {1.0f, 1.0f, 1.0f, 1.0f, 1.0f};
    int rc;
    for (uint32_t k = 0; k < nch; k++) {
        const uint32_t k0 = 5 * k, nk = std::min(5u, nterms - k0);
        if ((rc = lincomb5(c, nk, coefs + k0, d_terms + k0, sol ? d_solid_terms + k0 : nullptr, part[k],
                           sol ? spart[k] : nullptr)))
            return rc;
        tp[k] = part[k];
        sp[k] = sol ? spart[k] : nullptr;
    }
    return lincomb5(c, nch, ones, tp, sol ? sp : nullptr, d_out, sol ? d_solid_out : nullptr);
}

static int lincomb_check(arp_ctx* c, uint32_t nterms, const float* coefs, const ArpParticle* const* d_terms,
                         ArpParticle* d_out) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (nterms == 0 || !coefs || !d_terms || !d_out) return fail(c, ARP_E_ARG, "lincomb: bad arguments");
    for (uint32_t k = 0; k < nterms; k++)
        if (!d_terms[k]) return fail(c, ARP_E_ARG, "lincomb: term %u is null", k);
    if (nterms > 25) return fail(c, ARP_E_LIMIT, "lincomb of more than 25 terms");
    return ARP_OK;
}

int arp_lincomb(arp_ctx* c, uint32_t nterms, const float* coefs, const ArpParticle* const* d_terms,
                const ArpSolidParticle* const* d_solid_terms, ArpParticle* d_out, ArpSolidParticle* d_solid_out) {
    int rc;
    if ((rc = lincomb_check(c, nterms, coefs, d_terms, d_out))) return rc;
    if (nterms <= 5) return lincomb5(c, nterms, coefs, d_terms, d_solid_terms, d_out, d_solid_out);
    const bool sol = d_solid_terms && d_solid_out;
    const uint32_t nch = (nterms + 4) / 5;
    DBuf tmp[5], stmp[5];
    ArpParticle* part[5] = {};
    ArpSolidParticle* spart[5] = {};
    for (uint32_t k = 0; k < nch && rc == ARP_OK; k++) {
        rc = dalloc(c, tmp[k], sizeof(ArpParticle) * (size_t)std::max(c->n, 1u));
        if (rc == ARP_OK && sol) rc = dalloc(c, stmp[k], sizeof(ArpSolidParticle) * (size_t)std::max(c->ns, 1u));
        part[k] = P<ArpParticle>(tmp[k]);
        spart[k] = P<ArpSolidParticle>(stmp[k]);
    }
    if (rc == ARP_OK) rc = lincomb_chunked(c, nterms, coefs, d_terms, d_solid_terms, d_out, d_solid_out, part, spart);
    hipStreamSynchronize(c->stream);  // the partials are freed below
    for (uint32_t k = 0; k < nch; k++) {
        if (tmp[k].p) hipFree(tmp[k].p);
        if (stmp[k].p) hipFree(stmp[k].p);
    }
    return rc;
}

int arp_lincomb_noalloc(arp_ctx* c, uint32_t nterms, const float* coefs, const ArpParticle* const* d_terms,
                        const ArpSolidParticle* const* d_solid_terms, ArpParticle* d_out,
                        ArpSolidParticle* d_solid_out) {
    int rc;
    if ((rc = lincomb_check(c, nterms, coefs, d_terms, d_out))) return rc;
    if (nterms <= 5) return lincomb5(c, nterms, coefs, d_terms, d_solid_terms, d_out, d_solid_out);
    if (c->scratch_err || !c->scratch)
        return fail(c, ARP_E_HIP, "no device scratch on device %d (lookup error %d)", c->device, c->scratch_err);
    // the chunk partials live in this GPU's scratch array: particle partials, then solid ones,
    // each 256-byte aligned; every byte read is written first in the same call
    const bool sol = d_solid_terms && d_solid_out;
    const uint32_t nch = (nterms + 4) / 5;
    const auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t pb = al(sizeof(ArpParticle) * (size_t)c->n), sb = sol ? al(sizeof(ArpSolidParticle) * (size_t)c->ns) : 0;
    if (nch * (pb + sb) > c->scratch_bytes)
        return fail(c, ARP_E_LIMIT,
                    "lincomb_noalloc: %u terms of %u particles and %u solids need %zu bytes of partials, the "
                    "fixed device scratch holds %zu",
                    nterms, c->n, sol ? c->ns : 0, nch * (pb + sb), c->scratch_bytes);
    char* base = static_cast<char*>(c->scratch);
    ArpParticle* part[5] = {};
    ArpSolidParticle* spart[5] = {};
    for (uint32_t k = 0; k < nch; k++) {
        part[k] = reinterpret_cast<ArpParticle*>(base + k * pb);
        spart[k] = sol ? reinterpret_cast<ArpSolidParticle*>(base + nch * pb + k * sb) : nullptr;
    }
    return lincomb_chunked(c, nterms, coefs, d_terms, d_solid_terms, d_out, d_solid_out, part, spart);
}

int arp_velocity(arp_ctx* c, float r, const ArpParticle* d_in, const ArpSolidParticle* d_sin, ArpParticle* d_out,
                 ArpSolidParticle* d_sout) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (!d_in || !d_out) return fail(c, ARP_E_ARG, "velocity: null particle buffer");
    VelocityArgs a{c->n, (d_sin && d_sout) ? c->ns : 0, r, d_in, d_sin, d_out, d_sout};
    return launch(c, K_VELOCITY, nblk(std::max(c->n, a.ns)), 256, a);
}

int arp_set_neighbor_skin(arp_ctx* c, float skin, uint32_t cap, uint32_t flags) {
    if (!c) return ARP_E_ARG;
    if (!(skin >= 0.0f && skin <= 1.0f)) return fail(c, ARP_E_ARG, "skin must be in [0, 1] (a fraction of h)");
    if (skin > 0.0f && skin < 0.01f) return fail(c, ARP_E_ARG, "skin below 0.01 h leaves no room for rounding");
    if (cap == 0) cap = 48;
    if (cap < 8 || cap > 1024) return fail(c, ARP_E_ARG, "list capacity must be in [8, 1024]");
    c->nl_skin = skin;
    c->nl_cap = cap;
    if (!(c->nl_sskin < skin)) c->nl_sskin = 0.0f;  // a short skin must lie below the skin
    c->nl_always = (flags & ARP_NL_ALWAYS_BUILD) != 0;
    c->nl_nospec = (flags & ARP_NL_NO_SPECULATION) != 0;
    c->nl_dirty = true;
    // a replay must walk the recorded steps with the settings they were recorded with
    c->rec_valid = false;
    if (c->nlWords.p) {
        HIPCHK(c, hipSetDevice(c->device));
        FillArgs fa{P<uint32_t>(c->nlWords) + 8, 2, 0};  // statistics restart
        int rc = launch(c, K_FILL, 1, 256, fa);
        if (rc) return rc;
        FillArgs fr{P<uint32_t>(c->nlWords) + 18, 1, 0};
        if ((rc = launch(c, K_FILL, 1, 256, fr))) return rc;
    }
    return ARP_OK;
}

int arp_neighbor_lists_fit(uint32_t cap, uint32_t n) {
    if (cap == 0) cap = 48;
    return nl_fits_u32(cap, n) ? 1 : 0;
}

int arp_set_neighbor_max_age(arp_ctx* c, uint32_t stages) {
    if (!c) return ARP_E_ARG;
    c->nl_max_age = stages;
    c->nl_dirty = true;
    c->rec_valid = false;
    return ARP_OK;
}

int arp_neighbor_stats(arp_ctx* c, uint32_t* builds, uint32_t* reused) {
    if (!c) return ARP_E_ARG;
    uint32_t w[16] = {};
    if (c->nlWords.p) {
        HIPCHK(c, hipSetDevice(c->device));
        HIPCHK(c, hipMemcpyAsync(w, c->nlWords.p, sizeof w, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (builds) *builds = w[8];
    if (reused) *reused = w[9];
    return ARP_OK;
}

int arp_set_neighbor_short_skin(arp_ctx* c, float sskin) {
    if (!c) return ARP_E_ARG;
    if (!(sskin >= 0.0f && sskin <= 1.0f)) return fail(c, ARP_E_ARG, "short skin must be in [0, 1] (a fraction of h)");
    if (sskin > 0.0f && sskin < 0.01f) return fail(c, ARP_E_ARG, "short skin below 0.01 h leaves no room for rounding");
    if (sskin > 0.0f && !(sskin < c->nl_skin))
        return fail(c, ARP_E_ARG, "the short skin must lie below the list skin");
    c->nl_sskin = sskin;
    c->nl_dirty = true;
    c->rec_valid = false;  // a replay must walk the recorded steps with the lists they had
    if (c->nlWords.p) {
        HIPCHK(c, hipSetDevice(c->device));
        FillArgs fa{P<uint32_t>(c->nlWords) + 8, 2, 0};  // statistics restart
        int rc = launch(c, K_FILL, 1, 256, fa);
        if (rc) return rc;
        FillArgs fr{P<uint32_t>(c->nlWords) + 18, 1, 0};
        if ((rc = launch(c, K_FILL, 1, 256, fr))) return rc;
    }
    return ARP_OK;
}

int arp_neighbor_refreshes(arp_ctx* c, uint32_t* refreshes) {
    if (!c) return ARP_E_ARG;
    uint32_t w[kNlWords] = {};
    if (c->nlWords.p) {
        HIPCHK(c, hipSetDevice(c->device));
        HIPCHK(c, hipMemcpyAsync(w, c->nlWords.p, sizeof w, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (refreshes) *refreshes = w[18];
    return ARP_OK;
}

int arp_speculation_stats(const arp_ctx* c, uint64_t* steps, uint64_t* aborted) {
    if (!c) return ARP_E_ARG;
    if (steps) *steps = c->spec_steps;
    if (aborted) *aborted = c->spec_aborts;
    return ARP_OK;
}

int arp_profile(arp_ctx* c, int enable) {
    if (!c) return ARP_E_ARG;
    c->prof = enable != 0;
    return ARP_OK;
}

int arp_profile_select(arp_ctx* c, const char* kernels) {
    if (!c) return ARP_E_ARG;
    if (!kernels || !*kernels) {
        for (bool& b : c->prof_sel) b = true;
        return ARP_OK;
    }
    bool sel[K_COUNT] = {};
    std::string all(kernels);
    size_t pos = 0;
    while (pos <= all.size()) {
        size_t e = all.find(',', pos);
        if (e == std::string::npos) e = all.size();
        std::string name = all.substr(pos, e - pos);
        if (!name.empty()) {
            int k = 0;
            while (k < K_COUNT && name != kKernNames[k]) k++;
            if (k == K_COUNT) return fail(c, ARP_E_ARG, "unknown kernel '%s'", name.c_str());
            sel[k] = true;
        }
        pos = e + 1;
    }
    std::copy(sel, sel + K_COUNT, c->prof_sel);
    return ARP_OK;
}

int arp_profile_reset(arp_ctx* c) {
    if (!c) return ARP_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->ev_used = 0;
    c->ev_marks.clear();
    return ARP_OK;
}

int arp_profile_query(arp_ctx* c, const char* kernel, double* total_ms, uint64_t* launches) {
    if (!c || !kernel) return ARP_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    uint64_t cnt = 0;
    for (auto& m : c->ev_marks) {
        if (strcmp(kKernNames[m.first], kernel) != 0) continue;
        float ms = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev_pool[m.second], c->ev_pool[m.second + 1]));
        tot += ms;
        cnt++;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = cnt;
    return ARP_OK;
}

int arp_profile_durations(arp_ctx* c, const char* kernel, float* ms, uint64_t capacity, uint64_t* count) {
    if (!c || !kernel || (capacity && !ms)) return ARP_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    uint64_t cnt = 0;
    for (auto& m : c->ev_marks) {
        if (strcmp(kKernNames[m.first], kernel) != 0) continue;
        if (cnt < capacity) HIPCHK(c, hipEventElapsedTime(&ms[cnt], c->ev_pool[m.second], c->ev_pool[m.second + 1]));
        cnt++;
    }
    if (count) *count = cnt;
    return ARP_OK;
}

int arp_grid_info(arp_ctx* c, ArpGridInfo* out) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (!out) return ARP_E_ARG;
    memset(out, 0, sizeof *out);
    for (int i = 0; i < 3; i++) out->sub[i] = c->grid.sub[i];
    out->sub[3] = 1;
    out->ncells = c->ncells;
    out->n_valid[0] = c->bP_h.empty() ? 0 : c->bP_h.back();
    out->n_valid[1] = c->rP_h.empty() ? 0 : c->rP_h.back();
    if (!c->have_build) return ARP_OK;
    uint32_t nr = c->grid.nrows;
    std::vector<uint32_t> bs(nr + 2);
    uint32_t st[2] = {0, 0}, na = 0;
    HIPCHK(c, hipMemcpyAsync(bs.data(), c->binStart.p, 4 * (nr + 2), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&na, c->numActive.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(st, c->stats.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out->n_nonfinite[2] = bs[nr + 1] - bs[nr];
    out->n_overflow[2] = st[0];
    out->n_valid[2] = bs[nr] - st[0];
    out->occupied_rows = na;
    return ARP_OK;
}

int arp_neighbor_dump(arp_ctx* c, uint32_t* cell, uint32_t* cand_count, uint32_t* cand, uint64_t cap) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (!c->have_build) return fail(c, ARP_E_STATE, "no neighbour build to dump (call arp_neighbors_update)");
    const uint32_t n = c->n, nr = c->grid.nrows, sx = c->grid.sub[0], sy = c->grid.sub[1], sz = c->grid.sub[2];
    std::vector<uint32_t> perm(n), pid(n), bs(nr + 2), ro((size_t)nr * c->grid.rowstride);
    std::vector<float> pos(4 * (size_t)n);
    HIPCHK(c, hipMemcpyAsync(perm.data(), c->perm.p, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(pid.data(), c->last_pid->p, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(pos.data(), c->last_posd->p, 16 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(bs.data(), c->binStart.p, 4 * (nr + 2), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(ro.data(), c->rowOff.p, 4 * ro.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // sorted position of each particle (perm is the identity in a full build)
    std::vector<uint32_t> slot_of(n, ARP_INVALID_INDEX);
    for (uint32_t t = 0; t < n; t++) slot_of[pid[perm[t] & ~kFlag]] = t;
    std::vector<std::vector<uint32_t>> lists(n);
    for (uint32_t p = 0; p < n; p++) {
        uint32_t t = slot_of[p];
        uint32_t pe = perm[t];
        uint32_t i = pe & ~kFlag;
        if (pe & kFlag) {
            if (cell) cell[p] = ARP_INVALID_INDEX;
            continue;
        }
        uint32_t cx, cy, cz;
        host_cell(c->grid, &pos[4 * (size_t)i], cx, cy, cz);
        if (cell) cell[p] = cx + sx * (cy + sy * cz);
        uint32_t xl = cx > 0 ? cx - 1 : 0, xh = std::min(cx + 1, sx - 1);
        auto& L = lists[p];
        for (int dz = -1; dz <= 1; dz++) {
            int nz = (int)cz + dz;
            if (nz < 0 || nz >= (int)sz) continue;
            for (int dy = -1; dy <= 1; dy++) {
                int ny = (int)cy + dy;
                if (ny < 0 || ny >= (int)sy) continue;
                uint32_t row = (uint32_t)ny + sy * (uint32_t)nz;
                uint32_t s0 = bs[row], s1 = bs[row + 1];
                if (s1 > s0) {
                    const uint32_t* r = &ro[(size_t)row * c->grid.rowstride];
                    for (uint32_t q = r[xl]; q < r[xh + 1]; q++)
                        if (!(perm[q] & kFlag)) L.push_back(pid[perm[q] & ~kFlag]);
                }
                size_t cb = (size_t)row * sx;
                for (uint32_t q = c->bP_h[cb + xl]; q < c->bP_h[cb + xh + 1]; q++) L.push_back(c->b_orig[q] | kBnd);
            }
        }
        std::sort(L.begin(), L.end());
    }
    uint64_t off = 0;
    for (uint32_t p = 0; p < n; p++) {
        if (cand_count) cand_count[p] = (uint32_t)lists[p].size();
        for (uint32_t v : lists[p]) {
            if (cand && off < cap) cand[off] = v;
            off++;
        }
    }
    if (cand && off > cap) return fail(c, ARP_E_LIMIT, "candidate buffer too small (%llu needed)", (unsigned long long)off);
    return ARP_OK;
}

int arp_cell_counts(arp_ctx* c, uint64_t first, uint64_t count, uint32_t* out) {
    int rc;
    if ((rc = need_scene(c))) return rc;
    if (!out || first + count > c->ncells) return fail(c, ARP_E_ARG, "cell range out of bounds");
    const uint32_t nr = c->grid.nrows, sx = c->grid.sub[0];
    std::vector<uint32_t> bs(nr + 2), ro;
    if (c->have_build) {
        ro.resize((size_t)nr * c->grid.rowstride);
        HIPCHK(c, hipMemcpyAsync(bs.data(), c->binStart.p, 4 * (nr + 2), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(ro.data(), c->rowOff.p, 4 * ro.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    for (uint64_t k = 0; k < count; k++) {
        uint64_t cc = first + k;
        uint32_t b = c->bP_h[cc + 1] - c->bP_h[cc];
        uint32_t r = c->rP_h[cc + 1] - c->rP_h[cc];
        uint32_t p = 0;
        if (c->have_build) {
            uint32_t row = (uint32_t)(cc / sx), x = (uint32_t)(cc % sx);
            if (bs[row + 1] > bs[row]) {
                uint32_t m = ro[(size_t)row * c->grid.rowstride + x + 1] - ro[(size_t)row * c->grid.rowstride + x];
                p = std::min(m, kBucketCap - b - r);
            }
        }
        out[3 * k] = b;
        out[3 * k + 1] = r;
        out[3 * k + 2] = p;
    }
    return ARP_OK;
}


// ====================================================================================
// Batched ensembles: independent scenes (one context each) stepped together, one launch
// per kernel and RK stage for the whole batch, replayed from a captured hipGraph.
// ====================================================================================
}  // extern "C"

namespace {

// host-side per-step state of a context (what step_once changes besides device memory)
struct HostState {
    int cur;
    bool keys_valid, have_build, rec_valid;
    GridSlot slot;
};

HostState save_host(const arp_ctx* c) { return HostState{c->cur, c->keys_valid, c->have_build, c->rec_valid, c->slot}; }

void load_host(arp_ctx* c, const HostState& h) {
    c->cur = h.cur;
    c->keys_valid = h.keys_valid;
    c->have_build = h.have_build;
    c->rec_valid = h.rec_valid;
    c->slot = h.slot;
}

Kern batch_kernel(int k) {
    switch (k) {
        case K_KEY: return K_KEY_B;
        case K_SCAN: return K_SCAN_B;
        case K_SCATTER: return K_SCATTER_B;
        case K_ROWSORT: return K_ROWSORT_B;
        case K_FLUID: return K_FLUID_B;
        case K_FLUID_NL: return K_FLUID_NL_B;
        case K_FLUID_NLB: return K_FLUID_NLB_B;
        case K_FLUID_NL_F: return K_FLUID_NL_F_B;
        case K_STRAIN: return K_STRAIN_B;
        case K_SOLID: return K_SOLID_B;
        default: return K_COUNT;
    }
}

struct BLaunch {
    Kern k;
    uint32_t blocks, stride;
    DBuf args, scene, first;
    std::vector<uint8_t> blob;  // host copy (plan comparison)
    std::vector<uint32_t> hfirst, hscene;  // host sources of the uploads
};

}  // namespace

struct arp_batch {
    std::vector<arp_ctx*> ctx;
    arp_ctx* lead = nullptr;
    uint32_t flags = 0;
    std::vector<BLaunch> plan[2];
    std::vector<HostState> after[2];
    float plan_dt = 0.0f;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> ev;  // one per context + one for the batch stream
    std::string err;
    uint32_t walk_buf = 0;  // the list walks' LDS-buffered form (an autotuner choice)
};

namespace {

int bfail(arp_batch* b, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    b->err = buf;
    if (b->lead) b->lead->err = buf;
    return code;
}

void free_plan(arp_batch* b) {
    for (auto& P : b->plan)
        for (BLaunch& L : P)
            for (DBuf* d : {&L.args, &L.scene, &L.first}) free_buf(*d);
    b->plan[0].clear();
    b->plan[1].clear();
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Capture two consecutive steps of every scene (the device buffers alternate with period
// two) and lay them out as batched launches.  Host state ends as after two steps; the
// caller issues the steps and then sets the state for the number actually run.
int build_plan(arp_batch* b, float dt) {
    const size_t ns = b->ctx.size();
    arp_ctx* L0 = b->lead;
    int rc;
    for (arp_ctx* c : b->ctx)
        if (!c->keys_valid && !nl_enabled(c)) {  // bin the current state once so every step has the same launches
            if ((rc = run_key(c, P<float>(c->Yland[0])))) return rc;
            c->keys_valid = true;
        }
    std::vector<HostState> start(ns);
    for (size_t s = 0; s < ns; s++) start[s] = save_host(b->ctx[s]);
    // the fast-material walk when every scene has one mobile and one boundary material (each
    // scene's own, in its arguments), so that all scenes launch the same kernels
    bool allfm = true, anysolid = false;
    for (arp_ctx* c : b->ctx) {
        allfm = allfm && c->mobile_mat >= 0 && !c->any_elastic && !c->bnd_elastic && c->bnd_mat != -2;
        anysolid = anysolid || c->solids_active;
    }
    std::vector<BLaunch> plan[2];
    for (int p = 0; p < 2; p++) {
        std::vector<std::vector<Launch>> lists(ns);
        for (size_t s = 0; s < ns; s++) {
            arp_ctx* c = b->ctx[s];
            c->capture = &lists[s];
            c->capture_fm = allfm;
            c->capture_solid = anysolid;
            rc = step_once(c, dt);
            c->capture = nullptr;
            c->capture_fm = false;
            c->capture_solid = false;
            if (rc) {
                for (size_t t = 0; t < ns; t++) load_host(b->ctx[t], start[t]);
                return bfail(b, rc, "scene %zu: %s", s, c->err.c_str());
            }
            b->after[p].resize(ns);
            b->after[p][s] = save_host(c);
        }
        const size_t nl = lists[0].size();
        for (size_t s = 1; s < ns; s++) {
            bool same = lists[s].size() == nl;
            for (size_t j = 0; same && j < nl; j++) same = lists[s][j].k == lists[0][j].k;
            if (!same) {
                for (size_t t = 0; t < ns; t++) load_host(b->ctx[t], start[t]);
                return bfail(b, ARP_E_ARG, "scene %zu runs a different launch sequence (integrator or solid path differ)", s);
            }
        }
        for (size_t j = 0; j < nl; j++) {
            // a stage's list pair (build, then the fast-material walk with the same arguments)
            // becomes one launch of the walk's size, unless the walk is the buffered form
            bool pair = !b->walk_buf && j + 1 < nl && lists[0][j].k == K_FLUID_NLB && lists[0][j + 1].k == K_FLUID_NL_F;
            for (size_t s = 0; pair && s < ns; s++) pair = lists[s][j].args == lists[s][j + 1].args;
            if (pair) continue;  // (laid out with the walk, next)
            const bool fused = !b->walk_buf && j > 0 && lists[0][j - 1].k == K_FLUID_NLB && lists[0][j].k == K_FLUID_NL_F &&
                               std::all_of(lists.begin(), lists.end(),
                                           [&](const std::vector<Launch>& l) { return l[j - 1].args == l[j].args; });
            uint32_t total = 0;
            for (size_t s = 0; s < ns; s++) total += lists[s][j].blocks;
            if (total == 0) continue;  // no scene takes part
            BLaunch B;
            B.k = fused ? K_FLUID_NLBW_F_B : batch_kernel(lists[0][j].k);
            if (b->walk_buf && B.k == K_FLUID_NL_F_B) B.k = K_FLUID_NL_F_B_BUF;
            if (b->walk_buf && B.k == K_FLUID_NL_B) B.k = K_FLUID_NL_B_BUF;
            if (B.k == K_COUNT) {
                for (size_t t = 0; t < ns; t++) load_host(b->ctx[t], start[t]);
                return bfail(b, ARP_E_ARG, "kernel %s has no batched form (scenes with solids are not batched)",
                             kKernNames[lists[0][j].k]);
            }
            B.stride = (uint32_t)lists[0][j].args.size();
            std::vector<uint32_t> first(ns + 1, 0), scene;
            B.blob.resize((size_t)B.stride * ns);
            for (size_t s = 0; s < ns; s++) {
                const Launch& l = lists[s][j];
                memcpy(B.blob.data() + (size_t)s * B.stride, l.args.data(), B.stride);
                first[s + 1] = first[s] + l.blocks;
                scene.insert(scene.end(), l.blocks, (uint32_t)s);
            }
            B.blocks = first[ns];
            B.hfirst = std::move(first);
            B.hscene = std::move(scene);
            if (B.hscene.empty()) B.hscene.push_back(0);
            if ((rc = upload(L0, B.args, B.blob.data(), B.blob.size())) ||
                (rc = upload(L0, B.first, B.hfirst.data(), 4 * B.hfirst.size())) ||
                (rc = upload(L0, B.scene, B.hscene.data(), 4 * B.hscene.size())))
                return rc;
            plan[p].push_back(std::move(B));
        }
    }
    // keep the previous graph when the launches are unchanged
    bool same = b->plan[0].size() == plan[0].size() && b->plan[1].size() == plan[1].size();
    for (int p = 0; same && p < 2; p++)
        for (size_t j = 0; same && j < plan[p].size(); j++)
            same = plan[p][j].k == b->plan[p][j].k && plan[p][j].blob == b->plan[p][j].blob &&
                   plan[p][j].blocks == b->plan[p][j].blocks;
    if (same) {
        for (auto& P : plan)
            for (BLaunch& L : P)
                for (DBuf* d : {&L.args, &L.scene, &L.first}) free_buf(*d);
    } else {
        free_plan(b);
        b->plan[0] = std::move(plan[0]);
        b->plan[1] = std::move(plan[1]);
    }
    HIPCHK(L0, hipStreamSynchronize(L0->stream));  // the uploads read pageable host memory
    b->plan_dt = dt;
    return ARP_OK;
}

int issue(arp_batch* b, int p) {
    for (BLaunch& L : b->plan[p]) {
        BatchArgs a{L.args.p, L.stride, P<uint32_t>(L.scene), P<uint32_t>(L.first)};
        int rc = launch(b->lead, L.k, L.blocks, 256, a);
        if (rc) return rc;
    }
    return ARP_OK;
}

// order the batch stream after (before) the contexts' own streams
int join_streams(arp_batch* b, bool into_batch) {
    arp_ctx* L0 = b->lead;
    const size_t ns = b->ctx.size();
    if (b->ev.size() < ns + 1) {
        for (size_t i = b->ev.size(); i < ns + 1; i++) {
            hipEvent_t e;
            HIPCHK(L0, hipEventCreateWithFlags(&e, hipEventDisableTiming));
            b->ev.push_back(e);
        }
    }
    if (into_batch) {
        for (size_t s = 0; s < ns; s++) {
            arp_ctx* c = b->ctx[s];
            if (c->stream == L0->stream) continue;
            // (a context whose stream has finished its work needs no event)
            const hipError_t q = hipStreamQuery(c->stream);
            if (q == hipSuccess) continue;
            if (q != hipErrorNotReady) return bfail(b, ARP_E_HIP, "stream query failed: %s", hipGetErrorString(q));
            HIPCHK(L0, hipEventRecord(b->ev[s], c->stream));
            HIPCHK(L0, hipStreamWaitEvent(L0->stream, b->ev[s], 0));
        }
    } else {
        HIPCHK(L0, hipEventRecord(b->ev[ns], L0->stream));
        for (arp_ctx* c : b->ctx)
            if (c->stream != L0->stream) HIPCHK(L0, hipStreamWaitEvent(c->stream, b->ev[ns], 0));
    }
    return ARP_OK;
}

int run_batch(arp_batch* b, float dt, uint32_t nsteps) {
    arp_ctx* L0 = b->lead;
    int rc;
    if (nsteps == 0) return ARP_OK;
    // one launch sequence for all scenes: lists only when every scene keeps them
    bool all = true;
    for (arp_ctx* c : b->ctx) {
        c->nl_suspend = false;
        all = all && nl_enabled(c);
    }
    for (arp_ctx* c : b->ctx) c->nl_suspend = !all;
    struct Resume {
        arp_batch* b;
        ~Resume() {
            for (arp_ctx* c : b->ctx) c->nl_suspend = false;
        }
    } resume{b};
    for (arp_ctx* c : b->ctx) {  // outside the captured launches: reset stale list state
        if (nl_enabled(c) && (rc = nl_prepare(c))) return rc;
        if ((rc = sync_yland(c))) return rc;  // batched steps read and write the landing buffers
    }
    if ((rc = build_plan(b, dt))) return rc;
    if ((rc = join_streams(b, true))) return rc;
    const bool graph = (b->flags & ARP_BATCH_GRAPH) && !L0->prof && nsteps >= 2;
    if (graph) {
        if (!b->exec) {
            HIPCHK(L0, hipStreamBeginCapture(L0->stream, hipStreamCaptureModeThreadLocal));
            int rc2 = issue(b, 0);
            if (!rc2) rc2 = issue(b, 1);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(L0->stream, &g);
            if (rc2) return rc2;
            if (e != hipSuccess) return bfail(b, ARP_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
            b->graph = g;
            e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) return bfail(b, ARP_E_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
        }
        for (uint32_t k = 0; k < nsteps / 2; k++) HIPCHK(L0, hipGraphLaunch(b->exec, L0->stream));
        if (nsteps & 1u)
            if ((rc = issue(b, 0))) return rc;
    } else {
        for (uint32_t k = 0; k < nsteps; k++)
            if ((rc = issue(b, (int)(k & 1u)))) return rc;
    }
    const int last = (nsteps & 1u) ? 0 : 1;
    for (size_t s = 0; s < b->ctx.size(); s++) {
        load_host(b->ctx[s], b->after[last][s]);
        b->ctx[s]->time += (double)dt * nsteps;
    }
    return join_streams(b, false);
}

}  // namespace

extern "C" {

int arp_batch_create(arp_ctx* const* ctxs, uint32_t nscenes, uint32_t flags, arp_batch** out) {
    if (!out || !ctxs || nscenes == 0) return ARP_E_ARG;
    *out = nullptr;
    arp_ctx* L0 = ctxs[0];
    for (uint32_t s = 0; s < nscenes; s++) {
        arp_ctx* c = ctxs[s];
        if (!c) return ARP_E_ARG;
        if (!c->have_scene) return fail(L0, ARP_E_STATE, "scene %u has no scene uploaded", s);
        if (c->device != L0->device) return fail(L0, ARP_E_ARG, "scene %u is on another device", s);
        if (c->tab.stages == ARP_VERLET)
            return fail(L0, ARP_E_ARG, "scene %u uses velocity Verlet (batches run Runge-Kutta tableaus)", s);
        if (memcmp(&c->tab, &L0->tab, sizeof(ArpTableau)) != 0)
            return fail(L0, ARP_E_ARG, "scene %u uses another integrator", s);
        for (uint32_t t = 0; t < s; t++)
            if (ctxs[t] == c) return fail(L0, ARP_E_ARG, "scene %u appears twice", s);
    }
    arp_batch* b = new (std::nothrow) arp_batch();
    if (!b) return ARP_E_NOMEM;
    b->ctx.assign(ctxs, ctxs + nscenes);
    b->lead = L0;
    b->flags = flags;
    *out = b;
    return ARP_OK;
}

void arp_batch_destroy(arp_batch* b) {
    if (!b) return;
    hipSetDevice(b->lead->device);
    hipStreamSynchronize(b->lead->stream);
    free_plan(b);
    for (hipEvent_t e : b->ev) hipEventDestroy(e);
    delete b;
}

const char* arp_batch_last_error(const arp_batch* b) { return b ? b->err.c_str() : "null batch"; }

int arp_batch_step(arp_batch* b, float dt, uint32_t nsteps) {
    if (!b) return ARP_E_ARG;
    if (!(dt > 0.0f) || !std::isfinite(dt)) return bfail(b, ARP_E_ARG, "dt must be positive");
    HIPCHK(b->lead, hipSetDevice(b->lead->device));
    return run_batch(b, dt, nsteps);
}

int arp_batch_tuning(const arp_batch* b, uint32_t* rowsort_div, uint32_t* walk_buffered) {
    if (!b) return ARP_E_ARG;
    if (rowsort_div) *rowsort_div = b->lead->rowsort_div;
    if (walk_buffered) *walk_buffered = b->walk_buf;
    return ARP_OK;
}

int arp_batch_adjoint_record(arp_batch* b, float dt, uint32_t nsteps, uint32_t checkpoint_every) {
    if (!b) return ARP_E_ARG;
    for (size_t s = 0; s < b->ctx.size(); s++) {
        const int rc = arp_adjoint_record(b->ctx[s], dt, nsteps, checkpoint_every);
        if (rc) return bfail(b, rc, "scene %zu: %s", s, b->ctx[s]->err.c_str());
    }
    return ARP_OK;
}

int arp_batch_adjoint_backward(arp_batch* b, const void* const* d_grad_out, void* const* d_grad_in,
                               float* const* grad_params) {
    if (!b) return ARP_E_ARG;
    if (!d_grad_out || !d_grad_in) return bfail(b, ARP_E_ARG, "null cotangent array");
    for (size_t s = 0; s < b->ctx.size(); s++) {
        const int rc = arp_adjoint_backward(b->ctx[s], d_grad_out[s], d_grad_in[s], grad_params ? grad_params[s] : nullptr);
        if (rc) return bfail(b, rc, "scene %zu: %s", s, b->ctx[s]->err.c_str());
    }
    return ARP_OK;
}

int arp_batch_tune(arp_batch* b, float dt, uint32_t trial_steps, uint32_t* chosen_div) {
    if (!b) return ARP_E_ARG;
    arp_ctx* L0 = b->lead;
    HIPCHK(L0, hipSetDevice(L0->device));
    if (trial_steps == 0) trial_steps = 2;
    // states are saved and restored around every trial: tuning leaves the scenes untouched
    const size_t ns = b->ctx.size();
    std::vector<Checkpoint> ck(ns);
    std::vector<HostState> hs(ns);
    std::vector<double> tm(ns);
    int rc;
    for (size_t s = 0; s < ns; s++) {
        if ((rc = ck_save(b->ctx[s], ck[s]))) return rc;
        hs[s] = save_host(b->ctx[s]);
        tm[s] = b->ctx[s]->time;
    }
    // the knobs: particles per row-sort workgroup, and the list walks' form (registers only,
    // four waves per SIMD; or the entries within h buffered in LDS, five waves per SIMD)
    const uint32_t divs[] = {16, 32, 64, 128, 512, 2048};
    constexpr int kDiv = 6, kCand = 2 * kDiv;
    hipEvent_t e0, e1;
    HIPCHK(L0, hipEventCreate(&e0));
    HIPCHK(L0, hipEventCreate(&e1));
    // two rounds over the candidates, each keeping its faster trial (the first trials of a
    // process run measurably slower)
    float tbest[kCand];
    for (float& t : tbest) t = 1e30f;
    for (int round = 0; round < 2 && !rc; round++) {
        for (int ci = 0; ci < kCand; ci++) {
            const uint32_t div = divs[ci % kDiv];
            b->walk_buf = (uint32_t)(ci / kDiv);
            for (arp_ctx* c : b->ctx) c->rowsort_div = div;
            // two untimed steps build the plan, warm the caches and rebuild the lists that the
            // restore reset, so the timed steps see the reuse that follows (where a row sort
            // launched with many small workgroups still pays for every one of them)
            if ((rc = run_batch(b, dt, 2))) break;
            HIPCHK(L0, hipEventRecord(e0, L0->stream));
            if ((rc = run_batch(b, dt, trial_steps))) break;
            HIPCHK(L0, hipEventRecord(e1, L0->stream));
            HIPCHK(L0, hipEventSynchronize(e1));
            float ms = 0.0f;
            HIPCHK(L0, hipEventElapsedTime(&ms, e0, e1));
            tbest[ci] = std::min(tbest[ci], ms);
            for (size_t s = 0; s < ns; s++) {
                load_host(b->ctx[s], hs[s]);  // first: the restore writes the meta of this parity
                if ((rc = ck_restore(b->ctx[s], ck[s]))) break;
                b->ctx[s]->time = tm[s];
            }
            if (rc) break;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    uint32_t bestDiv = divs[0], bestWalk = 0;
    float best = 1e30f;
    for (int ci = 0; ci < kCand; ci++)
        if (tbest[ci] < best) {
            best = tbest[ci];
            bestDiv = divs[ci % kDiv];
            bestWalk = (uint32_t)(ci / kDiv);
        }
    for (arp_ctx* c : b->ctx) c->rowsort_div = bestDiv;
    b->walk_buf = bestWalk;
    for (Checkpoint& k : ck)
        for (DBuf* d : {&k.P, &k.V, &k.mk, &k.pid}) free_buf(*d);
    if (chosen_div) *chosen_div = bestDiv;
    return rc;
}

}  // extern "C"

namespace {

// (list skin, short skin) pairs the list tuner times, as fractions of h: one level at three
// skins, two levels at five
constexpr float kListCand[][2] = {{0.1f, 0.0f},  {0.2f, 0.0f},  {0.3f, 0.0f},  {0.2f, 0.04f},
                                  {0.3f, 0.03f}, {0.3f, 0.05f}, {0.4f, 0.04f}, {0.5f, 0.05f}};
constexpr int kNListCand = (int)(sizeof(kListCand) / sizeof(kListCand[0]));

int set_lists(arp_ctx* c, float skin, float sskin) {
    const uint32_t flags = (c->nl_always ? ARP_NL_ALWAYS_BUILD : 0u) | (c->nl_nospec ? ARP_NL_NO_SPECULATION : 0u);
    int rc = arp_set_neighbor_skin(c, skin, c->nl_cap, flags);
    if (!rc) rc = arp_set_neighbor_short_skin(c, sskin);
    return rc;
}

// Times trial steps of the contexts (as the batch b when given, else the one context) from
// their current states for every candidate pair, in two rounds keeping each candidate's
// faster trial, restoring the states after every trial; sets the fastest pair on every
// context.  Wall time around the whole call: it holds the list builds the pair causes and
// the host's share (the same for every pair).
int tune_lists(const std::vector<arp_ctx*>& cs, arp_batch* b, float dt, uint32_t trial, float* best_skin,
               float* best_sskin) {
    arp_ctx* L0 = cs[0];
    for (arp_ctx* c : cs)
        if (c->solids_active || c->n == 0)
            return fail(L0, ARP_E_STATE, "the list tuner needs fluid-only scenes with particles");
    if (!(dt > 0.0f) || !std::isfinite(dt)) return fail(L0, ARP_E_ARG, "dt must be positive");
    if (trial == 0) trial = 100;
    const size_t ns = cs.size();
    std::vector<Checkpoint> ck(ns);
    std::vector<HostState> hs(ns);
    std::vector<double> tm(ns);
    std::vector<uint64_t> sp(ns), sa(ns);
    int rc = ARP_OK;
    for (size_t s = 0; s < ns && !rc; s++) {
        rc = ck_save(cs[s], ck[s]);
        hs[s] = save_host(cs[s]);
        tm[s] = cs[s]->time;
        sp[s] = cs[s]->spec_steps;
        sa[s] = cs[s]->spec_aborts;
    }
    double tbest[kNListCand];
    for (double& t : tbest) t = 1e300;
    for (int round = 0; round < 2 && !rc; round++) {
        for (int ci = 0; ci < kNListCand && !rc; ci++) {
            for (arp_ctx* c : cs)
                if ((rc = set_lists(c, kListCand[ci][0], kListCand[ci][1]))) break;
            if (rc) break;
            for (arp_ctx* c : cs) HIPCHK(c, hipStreamSynchronize(c->stream));
            const auto t0 = std::chrono::steady_clock::now();
            rc = b ? run_batch(b, dt, trial) : run_steps(L0, dt, trial);
            for (arp_ctx* c : cs) HIPCHK(c, hipStreamSynchronize(c->stream));
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (!rc) tbest[ci] = std::min(tbest[ci], sec);
            for (size_t s = 0; s < ns && !rc; s++) {
                load_host(cs[s], hs[s]);  // first: the restore writes the meta of this parity
                rc = ck_restore(cs[s], ck[s]);
                cs[s]->time = tm[s];
            }
        }
    }
    int bi = 1;  // (the library's default pair when every trial failed)
    for (int ci = 0; ci < kNListCand; ci++)
        if (tbest[ci] < tbest[bi]) bi = ci;
    for (size_t s = 0; s < ns; s++) {
        int rc2 = set_lists(cs[s], kListCand[bi][0], kListCand[bi][1]);
        if (!rc) rc = rc2;
        cs[s]->spec_steps = sp[s];
        cs[s]->spec_aborts = sa[s];
    }
    for (Checkpoint& k : ck)
        for (DBuf* d : {&k.P, &k.V, &k.mk, &k.pid, &k.E}) free_buf(*d);
    if (best_skin) *best_skin = kListCand[bi][0];
    if (best_sskin) *best_sskin = kListCand[bi][1];
    return rc;
}

}  // namespace

extern "C" {

int arp_tune_lists(arp_ctx* c, float dt, uint32_t trial_steps, float* skin, float* short_skin) {
    int rc;
    if (!c) return ARP_E_ARG;
    if ((rc = need_scene(c))) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    return tune_lists(std::vector<arp_ctx*>{c}, nullptr, dt, trial_steps, skin, short_skin);
}

int arp_batch_tune_lists(arp_batch* b, float dt, uint32_t trial_steps, float* skin, float* short_skin) {
    if (!b) return ARP_E_ARG;
    HIPCHK(b->lead, hipSetDevice(b->lead->device));
    int rc = tune_lists(b->ctx, b, dt, trial_steps, skin, short_skin);
    if (rc) b->err = b->lead->err;
    return rc;
}

}  // extern "C"
