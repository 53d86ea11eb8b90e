// reseact_hip.cpp -- host side of libreseact_hip.so: the C ABI of
// include/reseact_hip.h on one MI355X (gfx950).
//
// Ownership: a context owns one HIP stream (default priority) and all device
// memory; every launch goes to that stream.  Adaptive solves never stop for
// the host per attempt: the PI controller runs on the device (rs_ctrl_update)
// and the host enqueues attempts in growing chunks, reading the controller
// block back once per chunk; kernels of a finished solve exit at their first
// instruction.  Chemistry kernels are compiled at rs_set_mechanism time with
// hipRTC from the mechanism compiler's straight-line code (mechgen.cpp).
//
// Reference interfaces replaced (EarthSciML/reseact.esm):
//   ssprk43_step / ros23_step / adaptive_solve / ssprk43_step_vjp /
//   ros23_step_vjp     tools/reactant_handoff/rx_traced_integrator.jl:279-623
//   macro_step, forward_pass, replay_fixed, backward_stage!, backward_sweep
//                      tools/adjoint_gradient.jl:538-834
//   lie_trotter_solve  tools/reactant_handoff/op_split.jl:100-164
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include "../../include/reseact_hip.h"
#include "mechgen.h"
#include "rs_launch.h"

// launch wrappers from transport.hip
extern "C" {
hipError_t rs_launch_transport(int kind, int tk, const RsTransArgs* a, hipStream_t st);
long long rs_transport_nblocks(int nlon, int nlat, int nlev, int nt, int tk);
int rs_transport_fused_plan(int nlon, int nlev, long long lds_bytes, int nth, int* plan);
long long rs_transport_fused_nblocks(int nlat, int nt, int nseg);
hipError_t rs_launch_transport_fused(int adj, const RsTransArgs* a, hipStream_t st);
hipError_t rs_launch_ctrl(RsCtrl* c, const double* partial, int npart, double* seqlog,
                          const double* negpart, hipStream_t st);
hipError_t rs_launch_fill(double* x, long long n, double v, hipStream_t st);
hipError_t rs_launch_fill_u64(unsigned long long* x, int n, unsigned long long v, hipStream_t st);
hipError_t rs_launch_mask(double* lam, const double* uout, long long n, hipStream_t st);
hipError_t rs_launch_clamp(double* x, long long n, hipStream_t st);
int rs_reduce_blocks(long long n);
hipError_t rs_launch_relmax(const double* a, const double* b, long long n, double* part, hipStream_t st);
hipError_t rs_launch_sumsq(const double* q, long long n, double* part, hipStream_t st);
hipError_t rs_launch_sum(const double* x, long long n, double* part, hipStream_t st);
hipError_t rs_launch_nonfinite(const double* x, long long n, unsigned long long* count,
                               hipStream_t st);
hipError_t rs_launch_dot(const double* w, const double* u, long long n, double* part, hipStream_t st);
hipError_t rs_launch_loss_seed(double* lam, const double* u, long long ntot, long long sbase,
                               long long ncol, hipStream_t st);
}
// from ppm.hip
long long rs_ppm_nblocks(int nlon, int nlat, int nlev, int nt);
hipError_t rs_launch_ppm_coefs(const double* dp, double* vs, double* ve, int nlon, int nlat,
                               int nlev, hipStream_t st);
hipError_t rs_launch_ppm_stage(int kind, const RsPPMArgs* a, hipStream_t st);
hipError_t rs_launch_ppm_pbl_setup(const double* dp, const double* pf, double* w, double* Wp,
                                   int nlon, int nlat, int nlev, hipStream_t st);
hipError_t rs_launch_ppm_vjp(const RsPPMVjpArgs* a, hipStream_t st);
hipError_t rs_launch_ppm_vjp_sweeps(const RsPPMVjpArgs* a, const double* F, double* G,
                                    double* divm, hipStream_t st);
hipError_t rs_launch_ppm_source(const double* Mx, const double* My, const double* dp,
                                const double* gx, const double* ce, const double* rdS,
                                const double* dPSdt, double* msrc, int nlon, int nlat, int nlev,
                                hipStream_t st);
hipError_t rs_launch_ppm_diag_mz(const double* Mx, const double* My, const double* dp,
                                 const double* gx, const double* ce, const double* rdS,
                                 const double* dB, const double* dPSdt, double* Mz, double* lid,
                                 int nlon, int nlat, int nlev, hipStream_t st);

// embedded sources of the runtime-compiled chemistry kernels
#include "jit_sources.inc"

#define KIND_STAGE 0
#define KIND_LAST 1
#define KIND_ADJ_A 2
#define KIND_ADJ_B 3
#define KIND_ADJ_C 4
#define KIND_ADJ_D 5

namespace {

enum ProfKind { P_TSTAGE, P_TLAST, P_CTRL, P_CHEM, P_PERCELL, P_CHEMVJP, P_TADJ, P_UTIL,
                P_TFUSED, P_TFUSEDADJ, P_TPPMADJ, P_CHEMUNEW, P_COUNT };
const char* kProfNames[P_COUNT] = {"transport_stage", "transport_last", "ctrl_update",
                                   "chem_attempt",    "chem_percell",   "chem_vjp",
                                   "transport_adj",   "util",           "transport_fused",
                                   "transport_fused_adj", "ppm_vjp",  "chem_unew"};

struct HipError {
    hipError_t e;
    const char* what;
};

#define HIPCHK(x)                                       \
    do {                                                \
        hipError_t e_ = (x);                            \
        if (e_ != hipSuccess) throw HipError{e_, #x};   \
    } while (0)

struct ArgError { std::string msg; int code; };

template <class T>
struct DevArray {
    T* p = nullptr;
    size_t n = 0;
    DevArray() = default;
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    ~DevArray() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void resize(size_t m) {
        if (m == n) return;
        release();
        if (m) HIPCHK(hipMalloc((void**)&p, m * sizeof(T)));
        n = m;
    }
    void ensure(size_t m) { if (m > n) resize(m); }
};

// The generated kernels come in RS_JIT_GROUPS code objects (compiled side by
// side), one module each.
#define RS_JIT_GROUPS 7
struct ChemModule {
    hipModule_t mod[RS_JIT_GROUPS] = {};
    hipFunction_t attempt = nullptr, attempt_kp = nullptr, attempt_pd = nullptr, percell = nullptr,
                  vjp = nullptr, vjp_mb = nullptr, percell_adj = nullptr;
    rs::MechPlan plan;
    ~ChemModule() {
        for (hipModule_t m : mod)
            if (m) (void)hipModuleUnload(m);
    }
};

// Device memory of rs_ros23_step_noalloc, for callers that cannot allocate:
// the error partials of one attempt (one per chemistry workgroup) and the
// fold of more than 1024 of them.  Every GPU the library runs on holds its
// own copy until the process ends; host code never passes `scratch` itself
// (a host-side stand-in), only the device address that hipGetSymbolAddress
// returns for the current device, kept per device in g_scratch_ptr.
#define RS_NOALLOC_FOLD 1024   /* RS_NOALLOC_PARTIALS: include/reseact_hip.h */
#define RS_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) double scratch[RS_NOALLOC_PARTIALS + RS_NOALLOC_FOLD];
std::mutex g_scratch_mu;
double* g_scratch_ptr[RS_NOALLOC_MAX_DEVICES] = {};

// Looks up this device's copy of `scratch` once (the device must be current).
void scratch_setup(int device) {
    if (device < 0 || device >= RS_NOALLOC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<double*>(ptr);
}

double* scratch_ptr(int device) {
    if (device < 0 || device >= RS_NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return g_scratch_ptr[device];
}

// Compiled code objects, shared by every context of the process.
std::mutex g_jit_mu;
std::map<std::string, std::vector<char>> g_jit_cache;

unsigned long long fnv1a(const std::string& s) {
    unsigned long long h = 1469598103934665603ULL;
    for (unsigned char c : s) { h ^= c; h *= 1099511628211ULL; }
    return h;
}

}  // namespace

struct rs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int ncu = 0;
    size_t lds_max = 65536;      // dynamic LDS a workgroup may take on this device
    size_t devmem = 0;           // device memory, bytes
    bool tape_stages = false;    // the replay tape keeps each PPM step's stage states
    std::vector<double> hK;      // rate constants (host copy: inline kernel argument)
    double* tape_view = nullptr; // replay tape lent by an ensemble (else `tape`)
    size_t tape_view_n = 0;
    std::vector<char> argbuf;    // chemistry kernel argument block
    int nlon = 0, nlat = 0, nlev = 0, ns = 0, nt = 0;
    long long nc = 0, ncol = 0;
    bool have_wind = false, have_mech = false, have_state = false, have_cand = false;
    int cur = 0;
    int lu_order = 1;
    int tk = 16;                 // transport tile height (levels per workgroup)
    int tmode = 1;               // 1: fused SSPRK43 kernels where the grid allows, 0: per stage
    int fnth = 1024;             // lanes per fused workgroup (1024 or 512)
    int fnth_req = 0;            // requested: 0 = by the grid's plan, else 1024 or 512
    int fplan[6] = {0, 0, 0, 0, 0, 0};  // fused plan: seg_w, seg_in, seg_halo, nseg, seg_lev, seg_chunks
    bool fplan_ok = false;
    int fplans[2][6] = {};       // the plans for 1024 and 512 lanes
    bool fplans_ok[2] = {false, false};
    unsigned chem_block = 64;    // chemistry threads per workgroup
    bool chem_kp = false;        // attempt kernel reads the rate constants through a pointer
    DevArray<double> S[2], X, Y, cx, wz, wl, K, jfac, E, lam, LA, LB, LC, gE, partial, negpart, seqlog;
    // monotone PPM transport (scheme RS_TRANSPORT_PPM)
    int scheme = RS_TRANSPORT_UPWIND3;
    int lon_periodic = 1;
    bool have_ppm_geom = false, have_ppm_flux = false;
    DevArray<double> pMx, pMy, pMz, pdp, pce, pgx, prdS, pvs, pve, pqw, pqe, pqs, pqn, pZ;
    DevArray<double> pF, pG, pdivm;   // PPM adjoint scratch: tendency, partial sums, div(M)
    DevArray<double> pdPS, pmsrc;   // pressure fixer: dPS/dt per column, air-mass source per cell
    bool have_fixer = false;
    // PBL mixing: layer fraction, weights dp*pbl_f, max(column weight, 1), stage means
    DevArray<double> ppblf, ppblw, ppblWp, ppblm;
    double pbl_tau = 0.0;
    bool have_pbl = false;
    DevArray<double> red;
    DevArray<unsigned long long> head;
    DevArray<unsigned> cmask;   // backward sweep: clamp masks as sign bits, two buffers
    // replayed chemistry attempts stop at unew (RsChemArgs.unew_only;
    // RS_REPLAY_UNEW=0: full attempts, for comparison)
    bool replay_unew = true;
    DevArray<long long> stats;
    DevArray<double> dtcell;
    RsCtrl* dctrl = nullptr;
    RsCtrl* hctrl = nullptr;       // pinned
    // Version of the state and the transport operator: every entry point
    // that may change S[cur], the operator or the transport scratch counts
    // it up.  A PPM step that keeps its stages records the version it left.
    unsigned long long sver = 0, kept_ver = ~0ull;
    bool keep_stages = false;
    double kept_dt = 0.0;
    long long* hstats = nullptr;   // pinned, 8
    double* hred = nullptr;        // pinned, 1024
    bool has_jfac = false, has_E = false;
    int nr = 0;
    std::unique_ptr<ChemModule> chem;
    // time-loop adjoint records
    std::vector<std::vector<std::pair<double, double>>> seqT, seqC;
    std::vector<std::pair<double, double>> lastseq[2];   // rs_get_step_log
    rs_split_opts last_opts{};
    bool have_record = false;
    DevArray<double> ckpt, tape, relpart;
    DevArray<unsigned long long> nfc;   // non-finite counts of the backward sweep's VJPs
    // per-cell chemistry in the last adjoint run: every macro step's seqC holds
    // one entry (t0, t1 - t0) for it; each cell started from rec_dtC0 and took
    // at most rec_pc_amax attempts
    bool rec_percell = false;
    double rec_dtC0 = 0.0;
    long long rec_pc_amax = 0;
    long long last_pc_amax = 0;         // attempts_max of the last per-cell solve
    DevArray<double> pcws, pcu;         // per-cell adjoint: lane records, state copy
    DevArray<unsigned long long> pcstats;
    DevArray<double> wobj;              // weights of the linear objective (loss kind 1)
    bool have_wobj = false;
    // profiling
    bool prof = false;
    struct EvPair { hipEvent_t a, b; int kind; };
    std::vector<EvPair> ev_pool;
    std::vector<hipEvent_t> ev_free;
    long long prof_n[P_COUNT] = {0};
    double prof_ms[P_COUNT] = {0};
    std::string err;

    size_t N() const { return (size_t)ns * (size_t)nc; }
};

namespace {

int fail(rs_ctx* c, int code, const std::string& m) {
    if (c) c->err = m;
    return code;
}

// ---- profiling --------------------------------------------------------------
void prof_flush(rs_ctx* c) {
    if (c->ev_pool.empty()) return;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& e : c->ev_pool) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
        c->prof_n[e.kind] += 1;
        c->prof_ms[e.kind] += ms;
        c->ev_free.push_back(e.a);
        c->ev_free.push_back(e.b);
    }
    c->ev_pool.clear();
}

hipEvent_t ev_get(rs_ctx* c) {
    if (c->ev_free.empty()) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        return e;
    }
    hipEvent_t e = c->ev_free.back();
    c->ev_free.pop_back();
    return e;
}

struct ProfScope {
    rs_ctx* c;
    hipEvent_t b{};
    int kind;
    bool on;
    ProfScope(rs_ctx* cc, int k) : c(cc), kind(k), on(cc->prof) {
        if (!on) return;
        if (c->ev_pool.size() >= 4096) prof_flush(c);
        hipEvent_t a = ev_get(c);
        b = ev_get(c);
        HIPCHK(hipEventRecord(a, c->stream));
        c->ev_pool.push_back({a, b, kind});
    }
    ~ProfScope() noexcept(false) {
        if (on) HIPCHK(hipEventRecord(b, c->stream));
    }
};

// ---- JIT --------------------------------------------------------------------
std::string chem_source(const rs::MechSpec& m, const rs::MechPlan& p) {
    std::string src;
    src += kJitPrelude;
    src += kJitLaunch;
    src += rs::generate_mech_source(m, p);
    src += kJitChemKernels;
    return src;
}

// On-disk cache of compiled code objects: $RS_JIT_CACHE_DIR, else
// $TMPDIR/reseact_jit_<uid>; RS_JIT_CACHE_DIR="" disables it.
std::string jit_cache_path(const std::string& key) {
    const char* env = std::getenv("RS_JIT_CACHE_DIR");
    std::string dir;
    if (env) {
        if (!*env) return "";
        dir = env;
    } else {
        const char* tmp = std::getenv("TMPDIR");
        dir = std::string(tmp && *tmp ? tmp : "/tmp") + "/reseact_jit_" + std::to_string(getuid());
    }
    mkdir(dir.c_str(), 0700);
    return dir + "/" + key + ".hsaco";
}

// Extra hipRTC options from RS_JIT_OPTIONS (space separated; developer knob
// for kernel variants, e.g. -DRS_PERCELL_WAVES=2); part of the cache key.
std::vector<std::string> jit_extra_options() {
    std::vector<std::string> out;
    if (const char* ev = std::getenv("RS_JIT_OPTIONS")) {
        std::istringstream is(ev);
        std::string w;
        while (is >> w) out.push_back(w);
    }
    return out;
}

std::vector<char> jit_compile(const std::string& src, bool exact, std::string& log,
                              const std::string& define = "") {
    std::vector<std::string> extra = jit_extra_options();
    if (!define.empty()) extra.push_back(define);
    std::string xk;
    for (auto& w : extra) xk += w + " ";
    const std::string key = std::to_string(fnv1a(src + "|" + xk)) + "_" + std::to_string(src.size()) +
                            (exact ? "_x" : "_f");
    {
        std::lock_guard<std::mutex> g(g_jit_mu);
        auto it = g_jit_cache.find(key);
        if (it != g_jit_cache.end()) return it->second;
    }
    const std::string path = jit_cache_path(key);
    if (!path.empty()) {
        if (FILE* f = std::fopen(path.c_str(), "rb")) {
            std::vector<char> code;
            char buf[65536];
            size_t n;
            while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) code.insert(code.end(), buf, buf + n);
            std::fclose(f);
            if (!code.empty()) {
                std::lock_guard<std::mutex> g(g_jit_mu);
                g_jit_cache[key] = code;
                return code;
            }
        }
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "rs_chem.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return {};
    }
    // exact (reference) mode: no FMA contraction and true division, so the
    // natural-order elimination rounds exactly like the reference's blocksolve
    std::vector<const char*> opts = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang",
                                     "-target-feature", "-Xclang", "-packed-fp32-ops"};
    if (exact) {
        opts.push_back("-ffp-contract=off");
        opts.push_back("-DRS_EXACT_DIV");
    }
    for (auto& w : extra) opts.push_back(w.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
    size_t ln = 0;
    hiprtcGetProgramLogSize(prog, &ln);
    if (ln > 1) {
        std::vector<char> lb(ln + 1, 0);
        hiprtcGetProgramLog(prog, lb.data());
        log = lb.data();
    }
    std::vector<char> code;
    if (r == HIPRTC_SUCCESS) {
        size_t cs = 0;
        hiprtcGetCodeSize(prog, &cs);
        code.resize(cs);
        hiprtcGetCode(prog, code.data());
    }
    hiprtcDestroyProgram(&prog);
    if (!code.empty()) {
        std::lock_guard<std::mutex> g(g_jit_mu);
        g_jit_cache[key] = code;
        if (!path.empty()) {
            const std::string tmpp = path + ".tmp" + std::to_string(getpid());
            if (FILE* f = std::fopen(tmpp.c_str(), "wb")) {
                const bool ok = std::fwrite(code.data(), 1, code.size(), f) == code.size();
                std::fclose(f);
                if (ok) std::rename(tmpp.c_str(), path.c_str());
                else std::remove(tmpp.c_str());
            }
        }
    }
    return code;
}

// ---- helpers -----------------------------------------------------------------
void h2d(double* d, const double* h, size_t n, hipStream_t st) {
    HIPCHK(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, st));
}
void d2h(double* h, const double* d, size_t n, hipStream_t st) {
    HIPCHK(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
}
void d2d(double* d, const double* s, size_t n, hipStream_t st) {
    if (n) HIPCHK(hipMemcpyAsync(d, s, n * sizeof(double), hipMemcpyDeviceToDevice, st));
}

RsTransArgs trans_args(rs_ctx* c) {
    RsTransArgs a{};
    a.S0 = c->S[0].p;
    a.S1 = c->S[1].p;
    a.cx = c->cx.p;
    a.wz = c->wz.p;
    a.wl = c->wl.p;
    a.nc = c->nc;
    a.nlon = c->nlon;
    a.nlat = c->nlat;
    a.nlev = c->nlev;
    a.nt = c->nt;
    a.partial = c->partial.p;
    return a;
}

bool trans_ppm(const rs_ctx* c) { return c->scheme == RS_TRANSPORT_PPM; }
bool trans_fused(const rs_ctx* c) { return c->tmode == 1 && c->fplan_ok && !trans_ppm(c); }

long long trans_nblocks(rs_ctx* c) {
    if (trans_ppm(c)) return rs_ppm_nblocks(c->nlon, c->nlat, c->nlev, c->nt);
    if (trans_fused(c)) return rs_transport_fused_nblocks(c->nlat, c->nt, c->fplan[3]);
    return rs_transport_nblocks(c->nlon, c->nlat, c->nlev, c->nt, c->tk);
}

void set_fused_plan(RsTransArgs& a, const rs_ctx* c) {
    a.seg_w = c->fplan[0];
    a.seg_in = c->fplan[1];
    a.seg_halo = c->fplan[2];
    a.nseg = c->fplan[3];
    a.seg_lev = c->fplan[4];
    a.seg_chunks = c->fplan[5];
    a.fnth = c->fnth;
}

// Fused workgroup size (1024 or 512 lanes) and its plan for the grid; nth 0:
// 1024 where a whole periodic row fits one workgroup (no halo recomputed),
// otherwise 512 (two workgroups per CU on the segments: 8-9% faster on
// 360x180x72, while 1024 is 7% faster on the whole rows of 144x91x47).
void select_fused(rs_ctx* c, int nth) {
    if (nth == 0) nth = (c->fplans_ok[0] && c->fplans[0][2] == 0) || !c->fplans_ok[1] ? 1024 : 512;
    const int q = nth == 512 ? 1 : 0;
    c->fnth = q ? 512 : 1024;
    c->fplan_ok = c->fplans_ok[q];
    for (int i = 0; i < 6; ++i) c->fplan[i] = c->fplans[q][i];
}

// Clamp flags of every step entry point: 1 = max(u, 0) on the accepted state
// (clamp_nonneg), 2 = positive domain (a trial state with a negative entry is
// rejected and retried at half the step).
inline int clamp_nonneg_of(int flags) { return flags & 1; }
inline int posdom_of(int flags) { return (flags >> 1) & 1; }

// One SSPRK43 attempt: either from the state buffers under a controller (ctrl
// != null: dt/cur/done from the device), or between explicit buffers.
// One SSPRK43 attempt of the PPM transport: three stages X, Y, X (u3) and the
// last stage from u and u3; input and output must differ (neighbours read
// the input while the output is written).
RsPPMArgs ppm_args(rs_ctx* c, double dt) {
    RsPPMArgs a{};
    a.S0 = c->S[0].p;
    a.S1 = c->S[1].p;
    a.partial = c->partial.p;
    a.Mx = c->pMx.p; a.My = c->pMy.p; a.Mz = c->pMz.p;
    a.ce = c->pce.p; a.gx = c->pgx.p; a.rdS = c->prdS.p;
    a.vs = c->pvs.p; a.ve = c->pve.p;
    a.qw = c->pqw.p; a.qe = c->pqe.p; a.qs = c->pqs.p; a.qn = c->pqn.p;
    a.msrc = c->pmsrc.n ? c->pmsrc.p : nullptr;
    if (c->have_pbl) {
        a.pblf = c->ppblf.p; a.pblw = c->ppblw.p; a.pblWp = c->ppblWp.p; a.pblm = c->ppblm.p;
        a.pbl_tau = c->pbl_tau;
    }
    a.nc = c->nc;
    a.dt_static = dt;
    a.nlon = c->nlon; a.nlat = c->nlat; a.nlev = c->nlev; a.nt = c->nt;
    a.cur_static = c->cur;
    a.lon_periodic = c->lon_periodic;
    return a;
}

// Stage states of one PPM step as its VJP reads them: u1, u2, u4 = a4 and
// the tendencies at u, u1, u2, u4 (F: four transported blocks).  The
// context's own set is X, Y, Z and the F slots; the time-loop adjoint keeps
// one set per replayed step in the tape.
struct PPMStages {
    double *u1, *u2, *u4, *F;
};

PPMStages ppm_ctx_stages(rs_ctx* c) { return PPMStages{c->X.p, c->Y.p, c->pZ.p, c->pF.p}; }

// keep (not null): also write the stage states and tendencies into *keep
// (u3 goes to LA), so that a VJP of this step at the same state reuses them.
void enqueue_ssprk43_ppm(rs_ctx* c, const RsCtrl* ctrl, const double* uin, double* uout,
                         double dt, double atol, double rtol, int flags,
                         const PPMStages* keep = nullptr) {
    const size_t NTC = (size_t)c->nt * c->nc;
    auto F = [&](int q) { return keep ? keep->F + (size_t)q * NTC : nullptr; };
    double* s1 = keep ? keep->u1 : c->X.p;
    double* s2 = keep ? keep->u2 : c->Y.p;
    double* s4 = keep ? keep->u4 : c->pZ.p;
    RsPPMArgs a = ppm_args(c, dt);
    a.ctrl = ctrl;
    a.posdom = posdom_of(flags);
    a.negpart = a.posdom ? c->negpart.p : nullptr;
    a.atol = atol;
    a.rtol = rtol;
    const bool st = (uin == nullptr);
    a.a_is_state = st ? 1 : 0;
    a.a = uin;
    a.out_is_state = 0;
    a.out = s1;
    a.clamp = 0;
    a.outf = F(0);
    { ProfScope ps(c, P_TSTAGE); HIPCHK(rs_launch_ppm_stage(0, &a, c->stream)); }
    a.a_is_state = 0;
    a.a = s1;
    a.out = s2;
    a.outf = F(1);
    { ProfScope ps(c, P_TSTAGE); HIPCHK(rs_launch_ppm_stage(0, &a, c->stream)); }
    // third stage: u3 into X (LA when keeping u1) and a4 = 2/3 u + 1/3 u3
    double* u3 = keep ? c->LA.p : c->X.p;
    a.a = s2;
    a.out = u3;
    a.b_is_state = st ? 1 : 0;
    a.b = uin;
    a.out2 = s4;
    a.outf = F(2);
    { ProfScope ps(c, P_TSTAGE); HIPCHK(rs_launch_ppm_stage(2, &a, c->stream)); }
    // last stage: stencils on a4, error from u and u3
    a.a_is_state = st ? 1 : 0;
    a.a = uin;
    a.b_is_state = 0;
    a.b = u3;
    a.c4 = s4;
    a.out2 = nullptr;
    a.out_is_state = st ? 1 : 0;
    a.out = uout;
    a.clamp = clamp_nonneg_of(flags);
    a.outf = F(3);
    { ProfScope ps(c, P_TLAST); HIPCHK(rs_launch_ppm_stage(1, &a, c->stream)); }
}

void enqueue_ssprk43(rs_ctx* c, const RsCtrl* ctrl, const double* uin, double* uout, double dt,
                     double atol, double rtol, int flags) {
    if (trans_ppm(c)) {
        enqueue_ssprk43_ppm(c, ctrl, uin, uout, dt, atol, rtol, flags);
        return;
    }
    const int clamp = clamp_nonneg_of(flags);
    RsTransArgs a = trans_args(c);
    a.posdom = posdom_of(flags);
    a.negpart = a.posdom ? c->negpart.p : nullptr;
    a.ctrl = ctrl;
    a.dt_static = dt;
    a.cur_static = c->cur;
    a.atol = atol;
    a.rtol = rtol;
    const bool st = (uin == nullptr);
    if (trans_fused(c) && (st || uin != uout)) {
        // all four stages in one launch; a segmented launch reads halo columns
        // other workgroups write, so input and output must differ (they do:
        // S[cur] -> S[cur^1], or two explicit buffers)
        set_fused_plan(a, c);
        a.a_is_state = st ? 1 : 0;
        a.a = uin;
        a.out_is_state = st ? 1 : 0;
        a.out = uout;
        a.clamp = clamp;
        ProfScope ps(c, P_TFUSED);
        HIPCHK(rs_launch_transport_fused(0, &a, c->stream));
        return;
    }
    // stage 1: X = u + h/2 L u
    a.a_is_state = st ? 1 : 0;
    a.a = uin;
    a.out_is_state = 0;
    a.out = c->X.p;
    a.clamp = 0;
    { ProfScope ps(c, P_TSTAGE); HIPCHK(rs_launch_transport(KIND_STAGE, c->tk, &a, c->stream)); }
    a.a_is_state = 0;
    a.a = c->X.p;
    a.out = c->Y.p;
    { ProfScope ps(c, P_TSTAGE); HIPCHK(rs_launch_transport(KIND_STAGE, c->tk, &a, c->stream)); }
    a.a = c->Y.p;
    a.out = c->X.p;
    { ProfScope ps(c, P_TSTAGE); HIPCHK(rs_launch_transport(KIND_STAGE, c->tk, &a, c->stream)); }
    a.a_is_state = st ? 1 : 0;
    a.a = uin;
    a.b = c->X.p;
    a.out_is_state = st ? 1 : 0;
    a.out = uout;
    a.clamp = clamp;
    { ProfScope ps(c, P_TLAST); HIPCHK(rs_launch_transport(KIND_LAST, c->tk, &a, c->stream)); }
}

// Lanes per chemistry workgroup (one lane per cell).
unsigned chem_threads(const rs_ctx* c) { return c->chem_block; }
// LDS of every runtime-compiled chemistry kernel for a workgroup of `block` lanes
size_t chem_slab_bytes(const rs_ctx* c, unsigned block) {
    return (size_t)block * RS_SLOW_DOUBLES(c->ns) * sizeof(double);
}
bool chem_block_fits(const rs_ctx* c, unsigned block) { return chem_slab_bytes(c, block) <= c->lds_max; }

// Launch of a runtime-compiled chemistry kernel with the LDS slab its
// workgroup size needs.  The argument block is the
// kernel's argument struct followed, for mechanisms of at most
// RS_K_INLINE_MAX reactions, by the rate constants (the kernels' second
// parameter; both are 8-byte aligned).
void chem_launch(rs_ctx* c, hipFunction_t f, const void* argp, size_t argsz, unsigned grid,
                 unsigned block, int kind, bool k_arg = true) {
    const size_t lds = chem_slab_bytes(c, block);
    const bool kin = k_arg && c->nr > 0 && c->nr <= RS_K_INLINE_MAX;
    const size_t off = (argsz + 7) & ~(size_t)7;
    size_t total = kin ? off + (size_t)c->nr * sizeof(double) : argsz;
    c->argbuf.assign(total, 0);
    std::memcpy(c->argbuf.data(), argp, argsz);
    if (kin) std::memcpy(c->argbuf.data() + off, c->hK.data(), (size_t)c->nr * sizeof(double));
    void* cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, c->argbuf.data(), HIP_LAUNCH_PARAM_BUFFER_SIZE,
                   &total, HIP_LAUNCH_PARAM_END};
    ProfScope ps(c, kind);
    HIPCHK(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, (unsigned)lds, c->stream, nullptr,
                                 cfg));
}




unsigned chem_grid(rs_ctx* c) { return (unsigned)((c->nc + chem_threads(c) - 1) / chem_threads(c)); }


// unew_only (static dt only): the step's new state alone, for the
// fixed-sequence replay; the attempt kernel returns before f2, k3 and the
// error estimate (its partials are then zero).
void enqueue_ros23(rs_ctx* c, const RsCtrl* ctrl, double* S0, double* S1, int cur_static,
                   double dt, double atol, double rtol, int flags, bool unew_only = false) {
    const int clamp = clamp_nonneg_of(flags);
    RsChemArgs a{};
    a.posdom = posdom_of(flags);
    a.negpart = a.posdom ? c->negpart.p : nullptr;
    a.S0 = S0;
    a.S1 = S1;
    a.ctrl = ctrl;
    a.K = c->K.p;
    a.jfac = c->has_jfac ? c->jfac.p : nullptr;
    a.E = c->has_E ? c->E.p : nullptr;
    a.partial = c->partial.p;
    a.nc = c->nc;
    a.ncol = c->ncol;
    a.dt_static = dt;
    a.atol = atol;
    a.rtol = rtol;
    a.cur_static = cur_static;
    a.clamp = clamp;
    a.unew_only = (unew_only && !ctrl) ? 1 : 0;
    const int kind = a.unew_only ? P_CHEMUNEW : P_CHEM;
    if (a.posdom) chem_launch(c, c->chem->attempt_pd, &a, sizeof(a), chem_grid(c), chem_threads(c), kind, false);
    else if (c->chem_kp) chem_launch(c, c->chem->attempt_kp, &a, sizeof(a), chem_grid(c), chem_threads(c), kind, false);
    else chem_launch(c, c->chem->attempt, &a, sizeof(a), chem_grid(c), chem_threads(c), kind);
}

// Sum of n device partials: folded on the device to at most 1024 values first.
double host_sum(rs_ctx* c, const double* dpart, int n) {
    if (n > 1024) {
        c->red.ensure(1024);
        HIPCHK(rs_launch_sum(dpart, n, c->red.p, c->stream));
        dpart = c->red.p;
        n = rs_reduce_blocks(n);
    }
    HIPCHK(hipMemcpyAsync(c->hred, dpart, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += c->hred[i];
    return s;
}

// max of n device partials; a NaN partial makes the result NaN
double host_max(rs_ctx* c, const double* dpart, int n) {
    if (n > 1024) throw ArgError{"host_max: more than 1024 partials", RS_E_ARG};
    HIPCHK(hipMemcpyAsync(c->hred, dpart, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (int i = 0; i < n; ++i)
        if (!(c->hred[i] <= s)) s = c->hred[i];
    return s;
}

enum class Op { T, C };

// Global-norm adaptive solve on the state buffers (rx_traced_integrator.jl:455-507).
// The global-norm adaptive solve (rx_traced_integrator.jl:455-507) in three
// pieces so that one context or a graph-captured ensemble can drive it:
// adaptive_begin fills and uploads the controller block, adaptive_chunk
// enqueues `n` attempts + controller updates, adaptive_finish reads back the
// outcome.  Attempts enqueued after the controller set `done` exit at once.
struct AdaptiveRun {
    Op op;
    double atol, rtol;
    int clamp;
    long long maxiters;
    int start_cur;
    bool log;
};

AdaptiveRun adaptive_begin(rs_ctx* c, Op op, double t0, double t1, double dt0, const rs_pictrl& pc,
                           double atol, double rtol, int clamp, long long maxiters, bool log,
                           hipStream_t st) {
    RsCtrl h{};
    h.t = t0;
    h.t1 = t1;
    h.tlim = t1 - 1.0e-9;
    h.dt = dt0;
    h.qold = pc.qoldinit;
    h.nnorm = (double)c->ns * (double)c->nc;
    h.beta1 = pc.beta1;
    h.beta2 = pc.beta2;
    h.invqmin = 1.0 / pc.qmin;
    h.invqmax = 1.0 / pc.qmax;
    h.gamma = pc.gamma;
    h.qsmin = pc.qsteady_min;
    h.qsmax = pc.qsteady_max;
    h.qoldinit = pc.qoldinit;
    h.maxiters = maxiters;
    h.cur = c->cur;
    h.done = (!(t0 < h.tlim) || maxiters <= 0) ? 1 : 0;
    h.dtc = std::min(h.dt, t1 - t0);
    long long cap = log ? std::min<long long>(maxiters, 1 << 22) : 0;
    if (log) c->seqlog.ensure((size_t)(2 * std::max<long long>(cap, 1)));
    h.logcap = (int)cap;
    *c->hctrl = h;
    HIPCHK(hipMemcpyAsync(c->dctrl, c->hctrl, sizeof(RsCtrl), hipMemcpyHostToDevice, st));
    return AdaptiveRun{op, atol, rtol, clamp, maxiters, c->cur, log};
}

void adaptive_chunk(rs_ctx* c, const AdaptiveRun& ar, int n) {
    const int npart = (ar.op == Op::T) ? (int)trans_nblocks(c) : (int)chem_grid(c);
    double* logp = ar.log ? c->seqlog.p : nullptr;
    for (int q = 0; q < n; ++q) {
        if (ar.op == Op::T) enqueue_ssprk43(c, c->dctrl, nullptr, nullptr, 0.0, ar.atol, ar.rtol, ar.clamp);
        else enqueue_ros23(c, c->dctrl, c->S[0].p, c->S[1].p, 0, 0.0, ar.atol, ar.rtol, ar.clamp);
        ProfScope ps(c, P_CTRL);
        HIPCHK(rs_launch_ctrl(c->dctrl, c->partial.p, npart, logp,
                              posdom_of(ar.clamp) ? c->negpart.p : nullptr, c->stream));
    }
}

// Requires c->hctrl to hold the final controller block (read back by the caller).
int adaptive_finish(rs_ctx* c, const AdaptiveRun& ar, double* dt_io, rs_counts* out,
                    std::vector<std::pair<double, double>>* seq) {
    const RsCtrl r = *c->hctrl;
    c->cur = r.cur;
    if (ar.op == Op::T && c->nt < c->ns && c->cur != ar.start_cur) {
        // species nt..ns-1 are untouched by transport: carry them over
        const size_t off = (size_t)c->nt * c->nc, n = (size_t)(c->ns - c->nt) * c->nc;
        d2d(c->S[c->cur].p + off, c->S[ar.start_cur].p + off, n, c->stream);
    }
    if (seq) {
        seq->clear();
        if (r.nlog > 0) {
            std::vector<double> lg(2 * (size_t)r.nlog);
            d2h(lg.data(), c->seqlog.p, lg.size(), c->stream);
            for (int i = 0; i < r.nlog; ++i) seq->emplace_back(lg[2 * i], lg[2 * i + 1]);
        }
        if (r.nacc > r.nlog) return fail(c, RS_E_ARG, "accepted-step log overflow");
    }
    *dt_io = r.dt;
    if (out) {
        out->naccept = r.nacc;
        out->nreject = r.nrej;
        out->t_end = r.t;
        out->dt_end = r.dt;
    }
    if (r.iters >= ar.maxiters && r.t < r.tlim)
        return fail(c, RS_E_MAXITERS, "adaptive loop hit maxiters at t=" + std::to_string(r.t));
    return RS_OK;
}

int run_adaptive(rs_ctx* c, Op op, double t0, double t1, double* dt_io, const rs_pictrl& pc,
                 double atol, double rtol, int clamp, long long maxiters, rs_counts* out,
                 std::vector<std::pair<double, double>>* seq) {
    const AdaptiveRun ar = adaptive_begin(c, op, t0, t1, *dt_io, pc, atol, rtol, clamp, maxiters,
                                          seq != nullptr, c->stream);
    int chunk = 2;
    while (!c->hctrl->done) {
        adaptive_chunk(c, ar, chunk);
        HIPCHK(hipMemcpyAsync(c->hctrl, c->dctrl, sizeof(RsCtrl), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        // doubling, but no more than the steps left at the current step size
        // plus one (a rejection): launches after the controller is done only
        // return, and the last doubling used to overshoot by up to 63
        int next = std::min(chunk * 2, 64);
        const RsCtrl& h = *c->hctrl;
        if (!h.done && h.dt > 0.0) {
            const double left = (h.t1 - h.t) / h.dt;
            if (left < next - 1) next = std::max(2, (int)std::ceil(left) + 1);
        }
        chunk = next;
    }
    return adaptive_finish(c, ar, dt_io, out, seq);
}

// Lanes of a persistent chemistry launch: as many workgroups as fit on the
// device at once, no more than the cells need.
long long percell_grid(rs_ctx* c, hipFunction_t f) {
    int per_cu = 0;
    const unsigned thr = chem_threads(c);
    HIPCHK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, (int)thr,
                                                              chem_slab_bytes(c, thr)));
    per_cu = std::max(per_cu, 1);
    long long grid = (long long)per_cu * c->ncu;
    grid = std::min<long long>(grid, (c->nc + thr - 1) / thr);
    return std::max<long long>(grid, 1);
}

// The controller fields of a per-cell argument block.
template <class A>
void percell_ctrl(A& a, const rs_pictrl& pc) {
    a.beta1 = pc.beta1;
    a.beta2 = pc.beta2;
    a.invqmin = 1.0 / pc.qmin;
    a.invqmax = 1.0 / pc.qmax;
    a.gamma = pc.gamma;
    a.qsmin = pc.qsteady_min;
    a.qsmax = pc.qsteady_max;
    a.qoldinit = pc.qoldinit;
}

// The per-cell solve on the state u (in place), statistics into c->stats;
// no host synchronisation.
void enqueue_percell(rs_ctx* c, double* u, double t0, double t1, double dt0, double* dtcell,
                     const rs_pictrl& pc, double atol, double rtol, int clamp, long long maxiters) {
    RsPercellArgs a{};
    a.u = u;
    a.K = c->K.p;
    a.jfac = c->has_jfac ? c->jfac.p : nullptr;
    a.E = c->has_E ? c->E.p : nullptr;
    a.dtcell = dtcell;
    c->head.ensure(1);
    c->stats.ensure(8);
    a.head = c->head.p;
    a.stats = c->stats.p;
    a.nc = c->nc;
    a.ncol = c->ncol;
    a.maxiters = maxiters;
    a.t0 = t0;
    a.t1 = t1;
    a.dt0 = dt0;
    a.atol = atol;
    a.rtol = rtol;
    percell_ctrl(a, pc);
    a.clamp = clamp_nonneg_of(clamp);
    a.posdom = posdom_of(clamp);
    // queue head and statistics (min starts at the largest value)
    long long init[8] = {0, 0, 0, 0, LLONG_MAX, 0, 0, 0};
    std::memcpy(c->hstats, init, sizeof(init));
    HIPCHK(hipMemcpyAsync(c->stats.p, c->hstats, sizeof(init), hipMemcpyHostToDevice, c->stream));
    HIPCHK(rs_launch_fill_u64(c->head.p, 1, 0ULL, c->stream));
    const long long grid = percell_grid(c, c->chem->percell);
    chem_launch(c, c->chem->percell, &a, sizeof(a), (unsigned)grid, chem_threads(c), P_PERCELL);
}

int run_percell(rs_ctx* c, double t0, double t1, double dt0, double* dt_cells_host,
                const rs_pictrl& pc, double atol, double rtol, int clamp, long long maxiters,
                rs_percell_stats* out, double* u = nullptr) {
    double* dtcell = nullptr;
    if (dt_cells_host) {
        c->dtcell.ensure((size_t)c->nc);
        h2d(c->dtcell.p, dt_cells_host, (size_t)c->nc, c->stream);
        dtcell = c->dtcell.p;
    }
    enqueue_percell(c, u ? u : c->S[c->cur].p, t0, t1, dt0, dtcell, pc, atol, rtol, clamp,
                    maxiters);
    HIPCHK(hipMemcpyAsync(c->hstats, c->stats.p, 8 * sizeof(long long), hipMemcpyDeviceToHost,
                          c->stream));
    if (dt_cells_host) d2h(dt_cells_host, c->dtcell.p, (size_t)c->nc, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->last_pc_amax = c->hstats[3];
    if (out) {
        out->attempts_total = c->hstats[0];
        out->accepts_total = c->hstats[1];
        out->rejects_total = c->hstats[2];
        out->attempts_max = c->hstats[3];
        out->attempts_min = c->hstats[4] == LLONG_MAX ? 0 : c->hstats[4];
        out->n_failed = c->hstats[5];
    }
    if (c->hstats[5] > 0)
        return fail(c, RS_E_MAXITERS, std::to_string(c->hstats[5]) + " cells hit maxiters");
    return RS_OK;
}

// Lane records of the per-cell adjoint for cells of at most nmax recorded
// states (allocated here, before any graph capture).
long long percell_adj_lanes(rs_ctx* c) {
    return percell_grid(c, c->chem->percell_adj) * (long long)chem_threads(c);
}
void ensure_percell_record(rs_ctx* c, long long nmax) {
    const double need = (double)percell_adj_lanes(c) * (double)nmax * (double)(c->ns + 1) * 8.0;
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    if (need > 0.85 * (double)(fr + c->pcws.n * sizeof(double)))
        throw ArgError{"per-cell adjoint record needs " + std::to_string(need / (1 << 30)) +
                       " GiB of device memory (" + std::to_string(nmax) + " states per cell)",
                       RS_E_ARG};
    c->pcws.ensure((size_t)percell_adj_lanes(c) * (size_t)nmax * (size_t)(c->ns + 1));
}

// The adjoint of the per-cell solve from the state u: lam (c->lam) from
// dJ/du(t1) to dJ/du(t0), c->gE accumulated, each cell's end state into uend
// (or not), nine counters into stats (zeroed here) and the non-finite
// emission-gradient contributions into nfc[1] (when nfc is given); no host
// synchronisation.
void enqueue_percell_adj(rs_ctx* c, const double* u, double* uend, double t0, double t1,
                         double dt0, const rs_pictrl& pc, double atol, double rtol, int clamp,
                         long long maxiters, long long nmax, unsigned long long* stats,
                         unsigned long long* nfc = nullptr) {
    RsPercellAdjArgs a{};
    a.nfc = nfc;
    a.u = u;
    a.uend = uend;
    a.lam = c->lam.p;
    a.gE = c->has_E ? c->gE.p : nullptr;
    a.K = c->K.p;
    a.jfac = c->has_jfac ? c->jfac.p : nullptr;
    a.E = c->has_E ? c->E.p : nullptr;
    a.ws = c->pcws.p;
    c->head.ensure(1);
    a.head = c->head.p;
    a.stats = stats;
    a.nc = c->nc;
    a.ncol = c->ncol;
    a.maxiters = maxiters;
    a.nmax = nmax;
    const long long grid = percell_grid(c, c->chem->percell_adj);
    a.nlanes = grid * (long long)chem_threads(c);
    if ((size_t)a.nlanes * (size_t)nmax * (size_t)(c->ns + 1) > c->pcws.n)
        throw ArgError{"per-cell adjoint record not sized", RS_E_STATE};
    a.t0 = t0;
    a.t1 = t1;
    a.dt0 = dt0;
    a.atol = atol;
    a.rtol = rtol;
    percell_ctrl(a, pc);
    a.clamp = clamp_nonneg_of(clamp);
    a.posdom = posdom_of(clamp);
    HIPCHK(rs_launch_fill_u64(stats, 9, 0ULL, c->stream));
    HIPCHK(rs_launch_fill_u64(c->head.p, 1, 0ULL, c->stream));
    chem_launch(c, c->chem->percell_adj, &a, sizeof(a), (unsigned)grid, chem_threads(c),
                P_PERCELL);
}

int check_ready(rs_ctx* c, bool need_mech, bool need_wind) {
    if (!c) return RS_E_ARG;
    if (c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!c->have_state) return fail(c, RS_E_STATE, "state not set");
    if (need_mech && !c->have_mech) return fail(c, RS_E_STATE, "mechanism not set");
    if (need_wind && c->nt > 0) {
        if (trans_ppm(c)) {
            if (!c->have_ppm_geom || !c->have_ppm_flux)
                return fail(c, RS_E_STATE, "PPM geometry or mass fluxes not set");
        } else if (!c->have_wind) {
            return fail(c, RS_E_STATE, "wind not set");
        }
    }
    return RS_OK;
}

int macro(rs_ctx* c, double t0, double t1, double* dtT, double* dtC, const rs_split_opts& o,
          rs_counts* cT, rs_counts* cC, std::vector<std::pair<double, double>>* sT,
          std::vector<std::pair<double, double>>* sC) {
    int rc = RS_OK;
    if (c->nt > 0) {
        rc = run_adaptive(c, Op::T, t0, t1, dtT, o.ctrl_T, o.atol_T, o.rtol_T,
                          o.clamp_nonneg & RS_CLAMP_NONNEG,
                          o.maxiters, cT, sT);
        if (rc) return rc;
    }
    if (o.chem_percell) {
        rs_percell_stats st{};
        rc = run_percell(c, t0, t1, *dtC, nullptr, o.ctrl_C, o.atol_C, o.rtol_C, o.clamp_nonneg,
                         o.maxiters, &st);
        if (cC) { cC->naccept = st.accepts_total; cC->nreject = st.rejects_total; cC->t_end = t1; cC->dt_end = *dtC; }
        if (sC) sC->assign(1, {t0, t1 - t0});   // the whole per-cell solve, as one entry
        return rc;
    }
    return run_adaptive(c, Op::C, t0, t1, dtC, o.ctrl_C, o.atol_C, o.rtol_C, o.clamp_nonneg,
                        o.maxiters, cC, sC);
}

// lam <- SSPRK43 VJP (lam in c->lam; species >= nt untouched: identity)
// lmask / nfcount (fused kernel only): the clamp mask of the step output
// applied as lambda is read, and the count of non-finite values written.
void enqueue_ssprk43_vjp(rs_ctx* c, double dt, const double* lmask = nullptr,
                         unsigned long long* nfcount = nullptr) {
    RsTransArgs a = trans_args(c);
    a.dt_static = dt;
    a.a_is_state = 0;
    a.out_is_state = 0;
    if (trans_fused(c)) {
        set_fused_plan(a, c);
        a.lmask = lmask;
        a.nfcount = nfcount;
        a.a = c->lam.p;
        if (c->fplan[2] == 0) {
            // whole periodic rows: each workgroup reads its slab before it
            // writes, and no other workgroup touches it, so in place is safe
            a.out = c->lam.p;
            ProfScope ps(c, P_TFUSEDADJ);
            HIPCHK(rs_launch_transport_fused(1, &a, c->stream));
        } else {
            a.out = c->LA.p;
            { ProfScope ps(c, P_TFUSEDADJ); HIPCHK(rs_launch_transport_fused(1, &a, c->stream)); }
            d2d(c->lam.p, c->LA.p, (size_t)c->nt * c->nc, c->stream);
        }
        return;
    }
    a.a = c->lam.p;
    a.out = c->LA.p;
    { ProfScope ps(c, P_TADJ); HIPCHK(rs_launch_transport(KIND_ADJ_A, c->tk, &a, c->stream)); }
    a.a = c->LA.p;
    a.out = c->LB.p;
    { ProfScope ps(c, P_TADJ); HIPCHK(rs_launch_transport(KIND_ADJ_B, c->tk, &a, c->stream)); }
    a.a = c->LB.p;
    a.out = c->LC.p;
    { ProfScope ps(c, P_TADJ); HIPCHK(rs_launch_transport(KIND_ADJ_C, c->tk, &a, c->stream)); }
    a.a = c->LC.p;
    a.b = c->LA.p;
    a.out = c->lam.p;
    { ProfScope ps(c, P_TADJ); HIPCHK(rs_launch_transport(KIND_ADJ_D, c->tk, &a, c->stream)); }
}

// lam <- SSPRK43 VJP of the PPM transport at the step input u (species >= nt
// untouched).  The tendency is nonlinear, so the stage states are recomputed
// with the forward stage kernels (X = u1, Y = u2, LA = u3, Z = u4 = 2/3 u +
// 1/3 u3, the forward's bits), then reversed with J at each stage state:
//   LB = lam + h/2 J(u4)^T lam                  (l4)
//   LC = x + h/2 J(u2)^T x,   x = 1/3 LB        (l2)
//   LA = LC + h/2 J(u1)^T LC                    (l1)
//   lam = 2/3 LB + (LA + h/2 J(u)^T LA)
// The PPM VJP uses line sweeps (each face gradient once), or the per-cell
// gather on a periodic row shorter than a stencil.
bool ppm_vjp_sweeps(const rs_ctx* c) { return !(c->lon_periodic && c->nlon < 6); }

// Scratch of the PPM VJP sweeps: the tendencies at u, u1, u2 (written by the
// recompute stages) and u4, the partial sums, div(M).  Sized before a graph
// capture, so that the enqueue below allocates nothing there.
void ppm_vjp_scratch(rs_ctx* c) {
    if (!ppm_vjp_sweeps(c)) return;
    c->pF.ensure(4 * (size_t)c->nt * c->nc);
    c->pG.ensure((size_t)std::max(c->nt - 1, 1) * c->nc);
    c->pdivm.ensure((size_t)c->nc);
}

// kept (not null): the stages of this step at u, written by a step taken
// with keep (sweeps only); otherwise they are recomputed into the context's
// set.
void enqueue_ssprk43_vjp_ppm(rs_ctx* c, const double* u, double dt,
                             const PPMStages* kept = nullptr) {
    const bool sweeps = ppm_vjp_sweeps(c);
    const size_t NT = (size_t)c->nt * c->nc;
    ppm_vjp_scratch(c);
    const PPMStages sg = kept ? *kept : ppm_ctx_stages(c);
    const bool stages = kept == nullptr;
    auto Fslot = [&](int q) { return sg.F + (size_t)q * NT; };
    RsPPMArgs a = ppm_args(c, dt);
    if (stages) {
    a.a = u;
    a.out = c->X.p;
    a.outf = sweeps ? Fslot(0) : nullptr;
    { ProfScope ps(c, P_TSTAGE); HIPCHK(rs_launch_ppm_stage(0, &a, c->stream)); }
    a.a = c->X.p;
    a.out = c->Y.p;
    a.outf = sweeps ? Fslot(1) : nullptr;
    { ProfScope ps(c, P_TSTAGE); HIPCHK(rs_launch_ppm_stage(0, &a, c->stream)); }
    a.a = c->Y.p;
    a.out = c->LA.p;
    a.b = u;
    a.out2 = c->pZ.p;
    a.outf = sweeps ? Fslot(2) : nullptr;
    { ProfScope ps(c, P_TSTAGE); HIPCHK(rs_launch_ppm_stage(2, &a, c->stream)); }
    }
    RsPPMVjpArgs v{};
    v.Mx = c->pMx.p; v.My = c->pMy.p; v.Mz = c->pMz.p;
    v.ce = c->pce.p; v.gx = c->pgx.p; v.rdS = c->prdS.p;
    v.vs = c->pvs.p; v.ve = c->pve.p;
    v.qw = c->pqw.p; v.qe = c->pqe.p; v.qs = c->pqs.p; v.qn = c->pqn.p;
    v.nc = c->nc;
    v.h2 = dt / 2;
    v.nlon = c->nlon; v.nlat = c->nlat; v.nlev = c->nlev; v.nt = c->nt;
    v.lon_periodic = c->lon_periodic;
    if (c->have_pbl) {
        v.pblf = c->ppblf.p; v.pblw = c->ppblw.p; v.pblWp = c->ppblWp.p; v.pbl_tau = c->pbl_tau;
    }
    if (sweeps && stages) {
        // the tendency at u4 (tendency-only stage)
        RsPPMArgs tf = ppm_args(c, dt);
        tf.a = c->pZ.p;
        tf.out = Fslot(3);
        ProfScope ps(c, P_TSTAGE);
        HIPCHK(rs_launch_ppm_stage(3, &tf, c->stream));
    }
    auto apply = [&](const double* at, int fq, const double* in, double alpha, const double* b,
                     double beta, double* out) {
        v.at = at; v.a = in; v.alpha = alpha; v.b = b; v.beta = beta; v.out = out;
        ProfScope ps(c, P_TPPMADJ);
        if (sweeps) HIPCHK(rs_launch_ppm_vjp_sweeps(&v, Fslot(fq), c->pG.p, c->pdivm.p, c->stream));
        else HIPCHK(rs_launch_ppm_vjp(&v, c->stream));
    };
    apply(sg.u4, 3, c->lam.p, 1.0, nullptr, 0.0, c->LB.p);
    apply(sg.u2, 2, c->LB.p, 1.0 / 3.0, nullptr, 0.0, c->LC.p);
    apply(sg.u1, 1, c->LC.p, 1.0, nullptr, 0.0, c->LA.p);
    apply(u, 0, c->LA.p, 1.0, c->LB.p, 2.0 / 3.0, c->lam.p);
}

void enqueue_ros23_vjp(rs_ctx* c, const double* u, const double* uout, double dt,
                       unsigned long long* nfc = nullptr,
                       const unsigned* mask_in = nullptr, unsigned* mask_out = nullptr) {
    RsChemVjpArgs a{};
    a.nfc = nfc;
    a.mask_in = mask_in;
    a.mask_out = mask_out;
    a.u = u;
    a.uout = uout;
    a.lam = c->lam.p;
    a.gE = c->gE.p;
    a.K = c->K.p;
    a.jfac = c->has_jfac ? c->jfac.p : nullptr;
    a.E = c->has_E ? c->E.p : nullptr;
    a.nc = c->nc;
    a.ncol = c->ncol;
    a.dt = dt;
    if (mask_in && !c->chem->vjp_mb)
        throw ArgError{"sign-bit clamp masks need at most 32 species", RS_E_STATE};
    chem_launch(c, mask_in ? c->chem->vjp_mb : c->chem->vjp, &a, sizeof(a), chem_grid(c),
                chem_threads(c), P_CHEMVJP);
}

// The objective at state u: loss kind 0, J = sum over level-0 cells of
// u[species]^2 (the workloads' quadratic loss); kind 1, J = <w, u> with the
// weights of rs_set_objective_weights (the reference's linear objective,
// adjoint_gradient.jl:408-426).
double loss_on_device(rs_ctx* c, const double* u, int kind, int species) {
    c->red.ensure(1024);
    if (kind == 1) {
        const long long n = (long long)c->N();
        HIPCHK(rs_launch_dot(c->wobj.p, u, n, c->red.p, c->stream));
        return host_sum(c, c->red.p, rs_reduce_blocks(n));
    }
    const double* q = u + (size_t)species * c->nc;
    const int nb = rs_reduce_blocks(c->ncol);
    HIPCHK(rs_launch_sumsq(q, c->ncol, c->red.p, c->stream));
    return host_sum(c, c->red.p, nb);
}

int check_loss(rs_ctx* c, int kind, int species) {
    if (kind == 0) {
        if (species < 0 || species >= c->ns) return fail(c, RS_E_ARG, "bad loss species");
        return RS_OK;
    }
    if (kind == 1) {
        if (!c->have_wobj) return fail(c, RS_E_STATE, "objective weights not set");
        return RS_OK;
    }
    return fail(c, RS_E_ARG, "bad loss kind");
}

// The replay tape: the ensemble's pool slice when one is lent, else the
// context's own buffer, grown to `n` doubles.
double* tape_ptr(rs_ctx* c, size_t n) {
    if (c->tape_view) {
        if (n > c->tape_view_n) throw ArgError{"replay tape slice too small", RS_E_STATE};
        return c->tape_view;
    }
    c->tape.ensure(n);
    return c->tape.p;
}

// Fixed-sequence replay of one macro step from `u0` into the tape; tape slot j
// holds the input of step j (transport steps first, then chemistry), the last
// slot the end state.  Returns the number of slots written.
//
// Per-cell chemistry (c->rec_percell; sC is its one entry (t0, t1 - t0)):
// with pc_forward the per-cell solve runs from slot nT into the last slot;
// without, the last slot is left to the per-cell adjoint (which writes the
// cells' end states there).
// Doubles of the tape for a macro step of nT transport and nC chemistry
// steps: nT + nC + 1 state slots, then (stages) the stage set of every PPM
// transport step: u1, u2, u4 and the four tendencies, 7 transported blocks.
size_t tape_doubles(rs_ctx* c, size_t nT, size_t nC, bool stages) {
    size_t n = (nT + nC + 1) * c->N();
    if (stages) n += nT * 7 * (size_t)c->nt * (size_t)c->nc;
    return n;
}

// Stage set of PPM transport step j in a tape of nslots state slots.
PPMStages tape_stage_set(rs_ctx* c, double* tp, size_t nslots, size_t j) {
    const size_t NTC = (size_t)c->nt * (size_t)c->nc;
    double* b = tp + nslots * c->N() + j * 7 * NTC;
    return PPMStages{b, b + NTC, b + 2 * NTC, b + 3 * NTC};
}

// Whether the backward sweeps keep the replayed PPM steps' stage states (their
// VJPs then skip the recompute): PPM transport with the sweep adjoint, and the
// stage sets of the longest recorded macro step within a quarter of the
// device memory.  Decided where the tape is sized.
bool want_tape_stages(rs_ctx* c, int nmacro) {
    if (!(c->nt > 0 && trans_ppm(c) && ppm_vjp_sweeps(c))) return false;
    size_t mx = 0;
    for (int m = 0; m < nmacro; ++m) mx = std::max(mx, c->seqT[m].size());
    const double bytes = (double)mx * 7.0 * (double)c->nt * (double)c->nc * sizeof(double);
    return bytes <= 0.25 * (double)c->devmem;
}

size_t replay_into_tape(rs_ctx* c, const double* u0, const std::vector<std::pair<double, double>>& sT,
                        const std::vector<std::pair<double, double>>& sC, const rs_split_opts& o,
                        bool pc_forward = true, bool stages = false) {
    const size_t N = c->N();
    const size_t nslots = sT.size() + sC.size() + 1;
    double* tp = tape_ptr(c, tape_doubles(c, sT.size(), sC.size(), stages));
    d2d(tp, u0, N, c->stream);
    size_t j = 0;
    for (auto& p : sT) {
        double* in = tp + j * N;
        double* outp = tp + (j + 1) * N;
        if (stages) {
            const PPMStages sg = tape_stage_set(c, tp, nslots, j);
            enqueue_ssprk43_ppm(c, nullptr, in, outp, p.second, o.atol_T, o.rtol_T,
                                clamp_nonneg_of(o.clamp_nonneg), &sg);
        } else {
            enqueue_ssprk43(c, nullptr, in, outp, p.second, o.atol_T, o.rtol_T,
                            clamp_nonneg_of(o.clamp_nonneg));
        }
        if (c->nt < c->ns) {
            const size_t off = (size_t)c->nt * c->nc;
            d2d(outp + off, in + off, (size_t)(c->ns - c->nt) * c->nc, c->stream);
        }
        ++j;
    }
    if (c->rec_percell) {
        if (pc_forward && !sC.empty()) {
            d2d(tp + (j + 1) * N, tp + j * N, N, c->stream);
            enqueue_percell(c, tp + (j + 1) * N, sC[0].first, sC[0].first + sC[0].second,
                            c->rec_dtC0, nullptr, o.ctrl_C, o.atol_C, o.rtol_C, o.clamp_nonneg,
                            o.maxiters);
        }
        return nslots;
    }
    for (auto& p : sC) {
        enqueue_ros23(c, nullptr, tp + j * N, tp + (j + 1) * N, 0, p.second, o.atol_C, o.rtol_C,
                      clamp_nonneg_of(o.clamp_nonneg), c->replay_unew);
        ++j;
    }
    return nslots;
}

// Size the replay tape once for the longest recorded macro step (before any
// graph capture: no allocation may happen while a stream is captured).
size_t tape_bytes(rs_ctx* c, int nmacro) {
    c->tape_stages = want_tape_stages(c, nmacro);
    size_t mx = c->N();
    for (int m = 0; m < nmacro; ++m)
        mx = std::max(mx, tape_doubles(c, c->seqT[m].size(), c->seqC[m].size(), c->tape_stages));
    return mx * sizeof(double);
}

void size_tape(rs_ctx* c, int nmacro) {
    const size_t n = tape_bytes(c, nmacro) / sizeof(double);
    if (!c->tape_view) c->tape.ensure(n);
}

// lambda = dJ/du(t_end) from the final checkpoint: 2 u[species] on level-0
// cells (kind 0) or the weights (kind 1); emission gradient cleared.  The
// non-finite counters of the sweep are cleared (a fill kernel, so that the
// graph-captured ensemble records no memset) and slot 0 gets the seed's count.
void backward_seed(rs_ctx* c, int kind, int species, int nvjp) {
    const size_t N = c->N();
    const int nm = (int)c->seqT.size();
    if (kind == 1) d2d(c->lam.p, c->wobj.p, N, c->stream);
    else HIPCHK(rs_launch_loss_seed(c->lam.p, c->ckpt.p + (size_t)nm * N, (long long)N,
                                    (long long)species * c->nc, c->ncol, c->stream));
    HIPCHK(rs_launch_fill(c->gE.p, (long long)c->gE.n, 0.0, c->stream));
    HIPCHK(rs_launch_fill_u64(c->nfc.p, 2 * (nvjp + 1), 0ULL, c->stream));
    if (c->ns <= 32 && c->have_mech) c->cmask.ensure(2 * (size_t)c->nc);   // sign bits, before any capture
    HIPCHK(rs_launch_nonfinite(c->lam.p, (long long)N, c->nfc.p, c->stream));
}

// VJPs of the backward sweep, and the counter slots before macro step m's
// (slot 0 is the seed; macro steps are swept last first).
int sweep_vjps(const rs_ctx* c) {
    size_t n = 0;
    for (size_t m = 0; m < c->seqT.size(); ++m) n += c->seqT[m].size() + c->seqC[m].size();
    if (n > (size_t)(INT_MAX / 2 - 2)) throw ArgError{"too many steps for the adjoint", RS_E_ARG};
    return (int)n;
}
size_t sweep_slot(const rs_ctx* c, int m) {
    size_t k = 1;
    for (size_t q = (size_t)m + 1; q < c->seqT.size(); ++q) k += c->seqT[q].size() + c->seqC[q].size();
    return k;
}

// The first VJP of the sweep that made lambda or its gradient non-finite
// (adjoint_gradient.jl:763-779: lambda_out has more non-finite entries than
// lambda_in, or the step's gradient contribution has any), from the counters
// read back: cnt[2 i], cnt[2 i + 1] for slot i.  Fills the bad_* fields of S;
// returns false when the sweep stayed finite.
bool first_nonfinite(const rs_ctx* c, const std::vector<unsigned long long>& cnt,
                     rs_adjoint_stats& S) {
    S.bad_op = -1;
    S.bad_macro = -1;
    S.bad_step = S.bad_nsteps = -1;
    S.bad_t = S.bad_dt = 0.0;
    S.bad_lam_in = S.bad_lam_out = S.bad_grad = 0;
    unsigned long long prev = cnt[0];
    size_t i = 1;
    for (int m = (int)c->seqT.size() - 1; m >= 0; --m) {
        for (int op = 1; op >= 0; --op) {
            const auto& sq = op ? c->seqC[m] : c->seqT[m];
            for (size_t q = sq.size(); q-- > 0; ++i) {
                const unsigned long long nl = cnt[2 * i], ng = cnt[2 * i + 1];
                if (nl > prev || ng > 0) {
                    S.bad_op = op;
                    S.bad_macro = m;
                    S.bad_step = (long long)q;
                    S.bad_nsteps = (long long)sq.size();
                    S.bad_t = sq[q].first;
                    S.bad_dt = sq[q].second;
                    S.bad_lam_in = (long long)prev;
                    S.bad_lam_out = (long long)nl;
                    S.bad_grad = (long long)ng;
                    return true;
                }
                prev = nl;
            }
        }
    }
    return false;
}

std::string nonfinite_message(const rs_adjoint_stats& S) {
    char b[320];
    std::snprintf(b, sizeof(b),
                  "non-finite adjoint: first at %s step %lld of %lld in macro step %d (t=%.9g "
                  "dt=%.9g): lambda non-finite %lld -> %lld, gradient non-finite %lld",
                  S.bad_op ? "chemistry" : "transport", S.bad_step + 1, S.bad_nsteps,
                  S.bad_macro + 1, S.bad_t, S.bad_dt, S.bad_lam_in, S.bad_lam_out, S.bad_grad);
    return b;
}

// Backward work of macro step m, enqueued on c->stream without host
// synchronisation (so it can be captured into a graph): fixed-sequence replay
// from checkpoint m into the tape (replay_fixed, adjoint_gradient.jl:647-661),
// the replay's deviation from checkpoint m+1 into `relpart` (:802-808), then
// the chemistry VJPs, last step first, and the transport VJPs
// (backward_stage!, :731-784; order :821-822), each after the clamp mask of
// its own step.  Optional events bracket the replay.
void enqueue_backward_macro(rs_ctx* c, int m, const rs_split_opts& o, double* relpart,
                            hipEvent_t ev0, hipEvent_t ev1) {
    unsigned long long* slot = c->nfc.p + 2 * sweep_slot(c, m);
    const size_t N = c->N();
    const auto& sT = c->seqT[m];
    const auto& sC = c->seqC[m];
    if (ev0) HIPCHK(hipEventRecord(ev0, c->stream));
    const size_t ns = replay_into_tape(c, c->ckpt.p + (size_t)m * N, sT, sC, o, false,
                                       c->tape_stages);
    double* tp = tape_ptr(c, ns * N);
    const size_t nT = sT.size();
    if (c->rec_percell) {
        // per-cell chemistry: each cell replayed and swept by the per-cell
        // adjoint, which writes the end states for the replay check
        if (ev1) HIPCHK(hipEventRecord(ev1, c->stream));
        enqueue_percell_adj(c, tp + nT * N, tp + (nT + 1) * N, sC[0].first,
                            sC[0].first + sC[0].second, c->rec_dtC0, o.ctrl_C, o.atol_C, o.rtol_C,
                            o.clamp_nonneg, o.maxiters, c->rec_pc_amax + 1,
                            c->pcstats.p + 9 * (size_t)m, slot);
        HIPCHK(rs_launch_relmax(tp + (ns - 1) * N, c->ckpt.p + (size_t)(m + 1) * N, (long long)N,
                                relpart, c->stream));
        ProfScope ps(c, P_UTIL);
        HIPCHK(rs_launch_nonfinite(c->lam.p, (long long)N, slot, c->stream));
        slot += 2;
    } else {
        HIPCHK(rs_launch_relmax(tp + (ns - 1) * N, c->ckpt.p + (size_t)(m + 1) * N, (long long)N,
                                relpart, c->stream));
        if (ev1) HIPCHK(hipEventRecord(ev1, c->stream));
    }
    // The clamp mask of chemistry step q is 1[u_{q+1} > 0], and u_{q+1} is the
    // input of step q + 1, whose VJP runs just before: that VJP writes its
    // input's sign bits (one word per cell, two buffers used in turn) and step
    // q reads them instead of u_{q+1} itself.  The last step reads its output.
    // RS_VJP_SIGN_BITS=0 turns this off (every VJP reads its output), so the
    // two paths can be compared.
    const char* sb = std::getenv("RS_VJP_SIGN_BITS");
    const bool clampC = clamp_nonneg_of(o.clamp_nonneg) != 0;
    const bool bits = clampC && c->ns <= 32 && c->cmask.n >= 2 * (size_t)c->nc &&
                      !(sb && sb[0] == '0');
    for (size_t q = c->rec_percell ? 0 : sC.size(); q-- > 0; slot += 2) {
        const size_t ts = nT + q;
        const bool last = q + 1 == sC.size();
        const unsigned* mi = (bits && !last) ? c->cmask.p + ((q + 1) & 1) * c->nc : nullptr;
        unsigned* mo = (bits && q > 0) ? c->cmask.p + (q & 1) * c->nc : nullptr;
        enqueue_ros23_vjp(c, tp + ts * N, (clampC && !mi) ? tp + (ts + 1) * N : nullptr,
                          sC[q].second, slot, mi, mo);
    }
    // The fused upwind VJP applies the mask as it reads lambda and counts the
    // non-finite values it writes, when it writes every entry (all species
    // transported); otherwise a mask pass before and a count pass after.
    const bool fold = trans_fused(c) && !trans_ppm(c) && c->nt == c->ns;
    for (size_t q = nT; q-- > 0; slot += 2) {
        if (fold) {
            enqueue_ssprk43_vjp(c, sT[q].second,
                                clamp_nonneg_of(o.clamp_nonneg) ? tp + (q + 1) * N : nullptr, slot);
            continue;
        }
        if (clamp_nonneg_of(o.clamp_nonneg)) {
            ProfScope ps(c, P_UTIL);
            HIPCHK(rs_launch_mask(c->lam.p, tp + (q + 1) * N, (long long)N, c->stream));
        }
        if (trans_ppm(c)) {
            const PPMStages sg = tape_stage_set(c, tp, ns, q);
            enqueue_ssprk43_vjp_ppm(c, tp + q * N, sT[q].second, c->tape_stages ? &sg : nullptr);
        } else {
            enqueue_ssprk43_vjp(c, sT[q].second);
        }
        ProfScope ps(c, P_UTIL);
        HIPCHK(rs_launch_nonfinite(c->lam.p, (long long)N, slot, c->stream));
    }
}

// max over `nsets` groups of `n` device partials (each group <= 1024); *worst
// gets the group of the largest (a NaN wins)
double host_max_many(rs_ctx* c, const double* dpart, int nsets, int n, int* worst = nullptr) {
    double mx = 0.0;
    for (int s = 0; s < nsets; ++s) {
        const double v = host_max(c, dpart + (size_t)s * 1024, n);
        if (!(v <= mx)) {
            mx = v;
            if (worst) *worst = s;
        }
    }
    return mx;
}

// The replay check of the reference (adjoint_gradient.jl:802-808): the
// fixed-sequence replay of every macro step must land within 1e-6 (relative,
// entrywise) of the next checkpoint.
std::string replay_failure(double d, int m) {
    if (d < 1e-6) return "";
    char b[200];
    std::snprintf(b, sizeof(b),
                  "fixed-sequence replay of macro step %d lands %.3g relative away from the "
                  "checkpointed end state: the tape is not the forward trajectory", m + 1, d);
    return b;
}

}  // namespace

namespace {
int guarded_keep(rs_ctx* c, const std::function<int()>& f);

// Every entry point runs through guarded (or guarded_keep, for the ones that
// leave the state, the transport operator and its scratch alone).
int guarded(rs_ctx* c, const std::function<int()>& f) {
    if (c) ++c->sver;
    return guarded_keep(c, f);
}

int guarded_keep(rs_ctx* c, const std::function<int()>& f) {
    try {
        return f();
    } catch (const HipError& e) {
        return fail(c, RS_E_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
    } catch (const ArgError& e) {
        return fail(c, e.code, e.msg);
    } catch (const std::bad_alloc&) {
        return fail(c, RS_E_ARG, "host out of memory");
    }
}
}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

const char* rs_backend(void) { return "hip-gfx950"; }

int rs_create(rs_ctx** out, int device) {
    if (!out) return RS_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RS_E_HIP;
    if (device < 0 || device >= ndev) return RS_E_ARG;
    auto* c = new rs_ctx();
    c->device = device;
    if (const char* ev = std::getenv("RS_REPLAY_UNEW")) c->replay_unew = ev[0] != '0';
    int rc = guarded(c, [&]() {
        HIPCHK(hipSetDevice(device));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        c->ncu = prop.multiProcessorCount;
        c->lds_max = prop.sharedMemPerBlock;
        c->devmem = prop.totalGlobalMem;
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPCHK(hipMalloc((void**)&c->dctrl, sizeof(RsCtrl)));
        HIPCHK(hipHostMalloc((void**)&c->hctrl, sizeof(RsCtrl), hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void**)&c->hstats, 8 * sizeof(long long), hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void**)&c->hred, 1024 * sizeof(double), hipHostMallocDefault));
        std::memset(c->hctrl, 0, sizeof(RsCtrl));
        scratch_setup(device);
        return RS_OK;
    });
    if (rc) {
        rs_destroy(c);
        return rc;
    }
    *out = c;
    return RS_OK;
}

int rs_destroy(rs_ctx* c) {
    if (!c) return RS_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& e : c->ev_pool) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    for (auto& e : c->ev_free) (void)hipEventDestroy(e);
    c->chem.reset();
    if (c->dctrl) (void)hipFree(c->dctrl);
    if (c->hctrl) (void)hipHostFree(c->hctrl);
    if (c->hstats) (void)hipHostFree(c->hstats);
    if (c->hred) (void)hipHostFree(c->hred);
    hipStream_t st = c->stream;
    delete c;   // DevArray destructors free device memory
    if (st) (void)hipStreamDestroy(st);
    return RS_OK;
}

const char* rs_last_error(const rs_ctx* c) { return c ? c->err.c_str() : "null context"; }

int rs_set_grid(rs_ctx* c, int nlon, int nlat, int nlev, int ns, int nt) {
    if (!c) return RS_E_ARG;
    if (nlon < 1 || nlat < 1 || nlev < 1 || ns < 1 || nt < 0 || nt > ns)
        return fail(c, RS_E_ARG, "bad grid sizes");
    if (nt > 0 && nlon < 4)
        return fail(c, RS_E_ARG, "transport needs nlon >= 4 (periodic 5-point stencil)");
    return guarded(c, [&]() {
        HIPCHK(hipSetDevice(c->device));
        c->nlon = nlon; c->nlat = nlat; c->nlev = nlev; c->ns = ns; c->nt = nt;
        c->ncol = (long long)nlon * nlat;
        c->nc = c->ncol * nlev;
        const size_t N = c->N(), NT = (size_t)nt * (size_t)c->nc;
        c->S[0].resize(N); c->S[1].resize(N);
        c->X.resize(NT); c->Y.resize(NT);
        c->lam.resize(N);
        c->LA.resize(NT); c->LB.resize(NT); c->LC.resize(NT);
        c->gE.resize((size_t)ns * c->ncol);
        long long np = std::max<long long>(rs_transport_nblocks(nlon, nlat, nlev, std::max(nt, 1), 4),
                                           (c->nc + 63) / 64);
        for (int q = 0; q < 2; ++q) {
            c->fplans_ok[q] = rs_transport_fused_plan(nlon, nlev, (long long)c->lds_max,
                                                      q ? 512 : 1024, c->fplans[q]) != 0;
            if (c->fplans_ok[q])
                np = std::max<long long>(np, rs_transport_fused_nblocks(nlat, std::max(nt, 1), c->fplans[q][3]));
        }
        select_fused(c, c->fnth_req);
        // PPM stages: one workgroup per tile of a level
        np = std::max<long long>(np, rs_ppm_nblocks(nlon, nlat, nlev, std::max(nt, 1)));
        c->partial.resize((size_t)np);
        c->negpart.resize((size_t)np);
        c->red.resize(1024);
        HIPCHK(rs_launch_fill(c->lam.p, (long long)N, 0.0, c->stream));
        HIPCHK(rs_launch_fill(c->gE.p, (long long)c->gE.n, 0.0, c->stream));
        c->cur = 0;
        c->scheme = RS_TRANSPORT_UPWIND3;
        c->pZ.resize(0);
        c->have_ppm_geom = c->have_ppm_flux = false;
        c->have_pbl = false;
        c->have_fixer = false;
        c->pmsrc.resize(0);
        {
            const size_t nq = nt > 1 ? (size_t)(nt - 1) : 0;
            const size_t nwe = nq * nlev * nlat, nsn = nq * nlev * nlon;
            c->pqw.resize(nwe); c->pqe.resize(nwe); c->pqs.resize(nsn); c->pqn.resize(nsn);
            if (nwe) {
                HIPCHK(rs_launch_fill(c->pqw.p, (long long)nwe, 0.0, c->stream));
                HIPCHK(rs_launch_fill(c->pqe.p, (long long)nwe, 0.0, c->stream));
            }
            if (nsn) {
                HIPCHK(rs_launch_fill(c->pqs.p, (long long)nsn, 0.0, c->stream));
                HIPCHK(rs_launch_fill(c->pqn.p, (long long)nsn, 0.0, c->stream));
            }
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        c->have_state = c->have_wind = c->have_mech = c->have_cand = false;
        c->has_jfac = c->has_E = false;
        c->chem.reset();
        c->have_record = false;
        c->have_wobj = false;
        return RS_OK;
    });
}

int rs_set_wind(rs_ctx* c, const double* cx, const double* wz, const double* wl) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!cx || !wz || !wl) return fail(c, RS_E_ARG, "null wind array");
    for (int j = 0; j < c->nlat; ++j) if (!std::isfinite(cx[j])) return fail(c, RS_E_NONFINITE, "cx");
    for (int k = 0; k <= c->nlev; ++k) if (!std::isfinite(wz[k])) return fail(c, RS_E_NONFINITE, "wz");
    for (int i = 0; i < c->nlon; ++i) if (!std::isfinite(wl[i])) return fail(c, RS_E_NONFINITE, "wl");
    return guarded(c, [&]() {
        c->cx.resize(c->nlat); c->wz.resize(c->nlev + 1); c->wl.resize(c->nlon);
        h2d(c->cx.p, cx, c->nlat, c->stream);
        h2d(c->wz.p, wz, c->nlev + 1, c->stream);
        h2d(c->wl.p, wl, c->nlon, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->have_wind = true;
        return RS_OK;
    });
}

int rs_set_transport_kernel(rs_ctx* c, int mode) {
    if (!c) return RS_E_ARG;
    if (mode < 0 || mode > 3) return fail(c, RS_E_ARG, "mode must be 0, 1, 2 or 3");
    c->tmode = mode == 0 ? 0 : 1;
    c->fnth_req = mode == 2 ? 1024 : (mode == 3 ? 512 : 0);
    select_fused(c, c->fnth_req);
    return RS_OK;
}

int rs_set_transport_scheme(rs_ctx* c, int scheme) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (scheme != RS_TRANSPORT_UPWIND3 && scheme != RS_TRANSPORT_PPM)
        return fail(c, RS_E_ARG, "unknown transport scheme");
    if (scheme == RS_TRANSPORT_PPM && c->nt < 1)
        return fail(c, RS_E_ARG, "PPM needs the air mass as transported species 0");
    return guarded(c, [&]() {
        // the PPM attempt keeps a4 of the third stage in a buffer of its own
        if (scheme == RS_TRANSPORT_PPM) c->pZ.resize((size_t)c->nt * (size_t)c->nc);
        else c->pZ.resize(0);
        c->scheme = scheme;
        return RS_OK;
    });
}

// Air-mass source of the column pressure fixer (recomputed with the fluxes
// and the geometry).
void ppm_update_source(rs_ctx* c) {
    if (!c->have_fixer || !c->have_ppm_geom || !c->have_ppm_flux) {
        c->pmsrc.resize(0);
        return;
    }
    c->pmsrc.resize(c->nc);
    HIPCHK(rs_launch_ppm_source(c->pMx.p, c->pMy.p, c->pdp.p, c->pgx.p, c->pce.p, c->prdS.p,
                                c->pdPS.p, c->pmsrc.p, c->nlon, c->nlat, c->nlev, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
}

// Weights of the PBL mixing (recomputed with dp) and the stage-mean scratch.
void ppm_update_pbl(rs_ctx* c) {
    if (!c->have_pbl || !c->have_ppm_flux) return;
    c->ppblw.resize(c->nc);
    c->ppblWp.resize(c->ncol);
    c->ppblm.resize((size_t)std::max(c->nt - 1, 1) * c->ncol);
    HIPCHK(rs_launch_ppm_pbl_setup(c->pdp.p, c->ppblf.p, c->ppblw.p, c->ppblWp.p, c->nlon, c->nlat,
                                   c->nlev, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
}

int rs_set_pbl_mixing(rs_ctx* c, const double* pbl_f, double tau_pblmix) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!pbl_f) {
        c->have_pbl = false;
        return RS_OK;
    }
    if (!c->have_ppm_flux) return fail(c, RS_E_STATE, "set the PPM mass fluxes (dp) first");
    if (!(tau_pblmix > 0.0) || !std::isfinite(tau_pblmix)) return fail(c, RS_E_ARG, "tau_pblmix must be positive");
    for (long long x = 0; x < c->nc; ++x)
        if (!std::isfinite(pbl_f[x])) return fail(c, RS_E_NONFINITE, "pbl_f");
    return guarded(c, [&]() {
        c->ppblf.resize(c->nc);
        h2d(c->ppblf.p, pbl_f, c->nc, c->stream);
        c->pbl_tau = tau_pblmix;
        c->have_pbl = true;
        ppm_update_pbl(c);
        return RS_OK;
    });
}

int rs_set_pressure_fixer(rs_ctx* c, const double* dPSdt) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!dPSdt) {
        c->have_fixer = false;
        return guarded(c, [&]() { ppm_update_source(c); return RS_OK; });
    }
    if (!c->have_ppm_geom || !c->have_ppm_flux)
        return fail(c, RS_E_STATE, "set the PPM geometry and mass fluxes first");
    for (long long x = 0; x < c->ncol; ++x)
        if (!std::isfinite(dPSdt[x])) return fail(c, RS_E_NONFINITE, "dPSdt");
    return guarded(c, [&]() {
        c->pdPS.resize(c->ncol);
        h2d(c->pdPS.p, dPSdt, c->ncol, c->stream);
        c->have_fixer = true;
        ppm_update_source(c);
        return RS_OK;
    });
}

int rs_diagnose_mz(rs_ctx* c, const double* dB, const double* dPSdt, double* Mz_out,
                   double* lid_resid) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!c->have_ppm_geom || !c->have_ppm_flux)
        return fail(c, RS_E_STATE, "set the PPM geometry and mass fluxes first");
    const int nlev = c->nlev;
    const long long ncol = c->ncol;
    if (dB)
        for (int k = 0; k < nlev; ++k)
            if (!std::isfinite(dB[k])) return fail(c, RS_E_NONFINITE, "dB");
    if (dPSdt)
        for (long long x = 0; x < ncol; ++x)
            if (!std::isfinite(dPSdt[x])) return fail(c, RS_E_NONFINITE, "dPSdt");
    return guarded(c, [&]() {
        DevArray<double> ddB, dps, dlid;
        if (dB) { ddB.resize(nlev); h2d(ddB.p, dB, nlev, c->stream); }
        if (dPSdt) { dps.resize(ncol); h2d(dps.p, dPSdt, ncol, c->stream); }
        dlid.resize(ncol);
        HIPCHK(rs_launch_ppm_diag_mz(c->pMx.p, c->pMy.p, c->pdp.p, c->pgx.p, c->pce.p, c->prdS.p,
                                     dB ? ddB.p : nullptr, dPSdt ? dps.p : nullptr, c->pMz.p,
                                     dlid.p, c->nlon, c->nlat, nlev, c->stream));
        std::vector<double> lid((size_t)ncol);
        HIPCHK(hipMemcpyAsync(lid.data(), dlid.p, sizeof(double) * ncol, hipMemcpyDeviceToHost,
                              c->stream));
        if (Mz_out)
            HIPCHK(hipMemcpyAsync(Mz_out, c->pMz.p, sizeof(double) * (size_t)(nlev + 1) * ncol,
                                  hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double worst = 0.0;
        for (double v : lid) {
            const double r = std::fabs(v);
            if (!(r <= worst)) worst = r;      // NaN-propagating
        }
        if (lid_resid) *lid_resid = worst;
        return RS_OK;
    });
}

int rs_set_ppm_geometry(rs_ctx* c, int lon_periodic, double dlon, const double* coslat_e,
                        const double* dS_lat, const double* dphi_lat) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!coslat_e || !dS_lat || !dphi_lat) return fail(c, RS_E_ARG, "null geometry array");
    if (!(dlon > 0.0) || !std::isfinite(dlon)) return fail(c, RS_E_ARG, "dlon must be positive");
    const int nlat = c->nlat;
    for (int j = 0; j <= nlat; ++j)
        if (!std::isfinite(coslat_e[j])) return fail(c, RS_E_NONFINITE, "geometry");
    for (int j = 0; j < nlat; ++j) {
        if (!std::isfinite(dphi_lat[j])) return fail(c, RS_E_NONFINITE, "geometry");
        if (!(dS_lat[j] > 0.0) || !std::isfinite(dS_lat[j])) return fail(c, RS_E_ARG, "dS_lat must be positive");
    }
    std::vector<double> gx(nlat), rdS(nlat);
    for (int j = 0; j < nlat; ++j) {
        gx[j] = dphi_lat[j] / (dlon * dS_lat[j]);
        rdS[j] = 1.0 / dS_lat[j];
    }
    return guarded(c, [&]() {
        c->pce.resize(nlat + 1); c->pgx.resize(nlat); c->prdS.resize(nlat);
        h2d(c->pce.p, coslat_e, nlat + 1, c->stream);
        h2d(c->pgx.p, gx.data(), nlat, c->stream);
        h2d(c->prdS.p, rdS.data(), nlat, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->lon_periodic = lon_periodic ? 1 : 0;
        c->have_ppm_geom = true;
        ppm_update_source(c);
        return RS_OK;
    });
}

int rs_set_mass_fluxes(rs_ctx* c, const double* Mx, const double* My, const double* Mz,
                       const double* dp) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!Mx || !My || !dp) return fail(c, RS_E_ARG, "null flux array");
    const size_t nlon = c->nlon, nlat = c->nlat, nlev = c->nlev, ncol = c->ncol;
    const size_t nx = nlev * nlat * (nlon + 1), ny = nlev * (nlat + 1) * nlon, nz = (nlev + 1) * ncol;
    for (size_t x = 0; x < nx; ++x) if (!std::isfinite(Mx[x])) return fail(c, RS_E_NONFINITE, "mass flux");
    for (size_t x = 0; x < ny; ++x) if (!std::isfinite(My[x])) return fail(c, RS_E_NONFINITE, "mass flux");
    if (Mz)
        for (size_t x = 0; x < nz; ++x) if (!std::isfinite(Mz[x])) return fail(c, RS_E_NONFINITE, "mass flux");
    for (size_t x = 0; x < (size_t)c->nc; ++x)
        if (!(dp[x] > 0.0) || !std::isfinite(dp[x])) return fail(c, RS_E_ARG, "dp must be positive");
    if (Mz)
        for (size_t x = 0; x < ncol; ++x)
            if (Mz[x] != 0.0 || Mz[nlev * ncol + x] != 0.0)
                return fail(c, RS_E_ARG, "Mz must vanish at the surface and the lid");
    return guarded(c, [&]() {
        c->pMx.resize(nx); c->pMy.resize(ny); c->pMz.resize(nz); c->pdp.resize(c->nc);
        c->pvs.resize(2 * (nlev + 4) * ncol); c->pve.resize(3 * (nlev + 3) * ncol);
        h2d(c->pMx.p, Mx, nx, c->stream);
        h2d(c->pMy.p, My, ny, c->stream);
        // Mz NULL: no vertical flux until rs_diagnose_mz supplies it
        if (Mz) h2d(c->pMz.p, Mz, nz, c->stream);
        else HIPCHK(rs_launch_fill(c->pMz.p, (long long)nz, 0.0, c->stream));
        h2d(c->pdp.p, dp, c->nc, c->stream);
        HIPCHK(rs_launch_ppm_coefs(c->pdp.p, c->pvs.p, c->pve.p, c->nlon, c->nlat, c->nlev,
                                   c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->have_ppm_flux = true;
        ppm_update_source(c);
        ppm_update_pbl(c);
        return RS_OK;
    });
}

int rs_set_inflow(rs_ctx* c, const double* qbc_w, const double* qbc_e, const double* qbc_s,
                  const double* qbc_n) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    const size_t nq = c->nt > 1 ? (size_t)(c->nt - 1) : 0;
    const size_t nwe = nq * c->nlev * c->nlat, nsn = nq * c->nlev * c->nlon;
    const double* src[4] = {qbc_w, qbc_e, qbc_s, qbc_n};
    const size_t n[4] = {nwe, nwe, nsn, nsn};
    for (int w = 0; w < 4; ++w)
        if (src[w])
            for (size_t x = 0; x < n[w]; ++x)
                if (!std::isfinite(src[w][x])) return fail(c, RS_E_NONFINITE, "inflow mixing ratio");
    return guarded(c, [&]() {
        DevArray<double>* dst[4] = {&c->pqw, &c->pqe, &c->pqs, &c->pqn};
        for (int w = 0; w < 4; ++w) {
            dst[w]->resize(n[w]);
            if (!n[w]) continue;
            if (src[w]) h2d(dst[w]->p, src[w], n[w], c->stream);
            else HIPCHK(rs_launch_fill(dst[w]->p, (long long)n[w], 0.0, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        return RS_OK;
    });
}

int rs_set_lu_ordering(rs_ctx* c, int ordering) {
    if (!c) return RS_E_ARG;
    if (ordering != 0 && ordering != 1) return fail(c, RS_E_ARG, "ordering must be 0 or 1");
    c->lu_order = ordering;
    return RS_OK;
}

}  // extern "C"

namespace {

// Compiles and loads the generated chemistry kernels of a validated spec.
int install_mechanism(rs_ctx* c, rs::MechSpec& m, const double* k) {
    const int nr = m.nr;
    for (int r = 0; r < nr; ++r)
        if (!std::isfinite(k[r])) return fail(c, RS_E_NONFINITE, "rate constant");
    std::string v = rs::validate_mech(m);
    if (!v.empty()) return fail(c, RS_E_ARG, v);
    return guarded(c, [&]() {
        auto cm = std::make_unique<ChemModule>();
        cm->plan = rs::plan_mechanism(m, c->lu_order);
        std::string log;
        // The exact-arithmetic build (no contraction, true division) is
        // refused: on gfx950 it has returned wrong values for some lanes on a
        // repeated launch, with no cause shown, so RS_JIT_EXACT=1 is an error
        // rather than a silently different result.
        if (const char* ev = std::getenv("RS_JIT_EXACT"))
            if (ev[0] == '1')
                return fail(c, RS_E_ARG, "RS_JIT_EXACT=1: the exact-arithmetic chemistry build "
                                         "is disabled (wrong lanes on repeated launches, cause "
                                         "not found)");
        // One program per kernel group (RS_JIT_ONLY), compiled on threads of
        // their own: a large mechanism's straight-line kernels take minutes
        // one after the other (RS_JIT_THREADS=0: in turn, on this thread).
        const std::string src = chem_source(m, cm->plan);
        std::vector<char> code[RS_JIT_GROUPS];
        std::string logs[RS_JIT_GROUPS];
        // the sign-bit VJP (group 7) exists for at most 32 species only
        const int ngroups = m.ns <= 32 ? RS_JIT_GROUPS : RS_JIT_GROUPS - 1;
        double secs[RS_JIT_GROUPS] = {};
        auto build = [&](int g) {
            const auto t0 = std::chrono::steady_clock::now();
            code[g] = jit_compile(src, false, logs[g], "-DRS_JIT_ONLY=" + std::to_string(g + 1));
            secs[g] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        };
        const char* jt = std::getenv("RS_JIT_THREADS");
        if (jt && jt[0] == '0') {
            for (int g = 0; g < ngroups; ++g) build(g);
        } else {
            std::vector<std::thread> th;
            for (int g = 0; g < ngroups; ++g) th.emplace_back(build, g);
            for (auto& t : th) t.join();
        }
        if (const char* tv = std::getenv("RS_JIT_TIMES"))   // developer knob: compile times
            if (tv[0] == '1')
                for (int g = 0; g < ngroups; ++g)
                    std::fprintf(stderr, "[reseact] chemistry kernel group %d: %.1f s\n", g + 1,
                                 secs[g]);
        for (int g = 0; g < ngroups; ++g)
            if (code[g].empty())
                return fail(c, RS_E_JIT, "hipRTC compile failed: " + logs[g].substr(0, 2000));
        HIPCHK(hipSetDevice(c->device));
        for (int g = 0; g < ngroups; ++g) HIPCHK(hipModuleLoadData(&cm->mod[g], code[g].data()));
        HIPCHK(hipModuleGetFunction(&cm->attempt, cm->mod[0], "rs_chem_attempt"));
        HIPCHK(hipModuleGetFunction(&cm->attempt_kp, cm->mod[1], "rs_chem_attempt_kp"));
        HIPCHK(hipModuleGetFunction(&cm->attempt_pd, cm->mod[2], "rs_chem_attempt_pd"));
        HIPCHK(hipModuleGetFunction(&cm->percell, cm->mod[3], "rs_chem_percell"));
        HIPCHK(hipModuleGetFunction(&cm->percell_adj, cm->mod[4], "rs_chem_percell_adj"));
        HIPCHK(hipModuleGetFunction(&cm->vjp, cm->mod[5], "rs_chem_vjp"));
        if (ngroups == RS_JIT_GROUPS)
            HIPCHK(hipModuleGetFunction(&cm->vjp_mb, cm->mod[6], "rs_chem_vjp_mb"));
        c->K.resize(std::max(nr, 1));
        if (nr) h2d(c->K.p, k, nr, c->stream);
        c->hK.assign(k, k + nr);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->nr = nr;
        c->chem = std::move(cm);
        c->have_mech = true;
        c->partial.ensure((size_t)chem_grid(c));   // one partial per chemistry workgroup
        c->negpart.ensure((size_t)chem_grid(c));
        return RS_OK;
    });
}

}  // namespace

extern "C" {

int rs_set_mechanism(rs_ctx* c, int nr, const int* re, const int* pr, const double* k,
                     const int* photo) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (nr < 0 || (nr > 0 && (!re || !pr || !k))) return fail(c, RS_E_ARG, "bad mechanism arrays");
    rs::MechSpec m;
    m.ns = c->ns;
    m.nr = nr;
    for (int r = 0; r < nr; ++r) {
        m.ra.push_back(re[2 * r]); m.rb.push_back(re[2 * r + 1]);
        m.pa.push_back(pr[2 * r]); m.pb.push_back(pr[2 * r + 1]);
        m.photo.push_back(photo ? (photo[r] != 0) : 0);
    }
    return install_mechanism(c, m, k);
}

int rs_set_mechanism_stoich(rs_ctx* c, int nr, const int* re, const int* prod_off,
                            const int* prod_species, const double* prod_coef, const double* k,
                            const int* photo) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (nr < 0 || (nr > 0 && (!re || !prod_off || !k))) return fail(c, RS_E_ARG, "bad mechanism arrays");
    if (nr > 0 && (prod_off[0] != 0 || (prod_off[nr] > 0 && (!prod_species || !prod_coef))))
        return fail(c, RS_E_ARG, "bad product offsets");
    rs::MechSpec m;
    m.ns = c->ns;
    m.nr = nr;
    m.prod.resize(nr);
    for (int r = 0; r < nr; ++r) {
        if (prod_off[r + 1] < prod_off[r]) return fail(c, RS_E_ARG, "bad product offsets");
        m.ra.push_back(re[2 * r]); m.rb.push_back(re[2 * r + 1]);
        m.pa.push_back(-1); m.pb.push_back(-1);
        m.photo.push_back(photo ? (photo[r] != 0) : 0);
        for (int q = prod_off[r]; q < prod_off[r + 1]; ++q)
            m.prod[r].emplace_back(prod_species[q], prod_coef[q]);
    }
    return install_mechanism(c, m, k);
}

int rs_set_photolysis(rs_ctx* c, const double* jfac) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    return guarded(c, [&]() {
        if (!jfac) { c->has_jfac = false; return RS_OK; }
        c->jfac.resize((size_t)c->nc);
        h2d(c->jfac.p, jfac, (size_t)c->nc, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->has_jfac = true;
        return RS_OK;
    });
}

int rs_set_emissions(rs_ctx* c, const double* e) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    return guarded(c, [&]() {
        if (!e) { c->has_E = false; return RS_OK; }
        // kept cell-major on the device, [ncol][ns]: a chemistry lane reads
        // its cell's ns values at fixed offsets from one address
        const size_t ns = (size_t)c->ns, ncol = (size_t)c->ncol;
        std::vector<double> cm(ns * ncol);
        for (size_t s = 0; s < ns; ++s)
            for (size_t q = 0; q < ncol; ++q) cm[q * ns + s] = e[s * ncol + q];
        c->E.resize(ns * ncol);
        h2d(c->E.p, cm.data(), c->E.n, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->has_E = true;
        return RS_OK;
    });
}

int rs_set_state(rs_ctx* c, const double* u) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!u) return fail(c, RS_E_ARG, "null state");
    return guarded(c, [&]() {
        h2d(c->S[c->cur].p, u, c->N(), c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->have_state = true;
        c->have_cand = false;
        return RS_OK;
    });
}

int rs_get_state(rs_ctx* c, double* u) {
    if (!c || !c->have_state) return fail(c, RS_E_STATE, "state not set");
    if (!u) return fail(c, RS_E_ARG, "null output");
    return guarded_keep(c, [&]() { d2h(u, c->S[c->cur].p, c->N(), c->stream); return RS_OK; });
}

int rs_clamp_state(rs_ctx* c) {
    if (!c || !c->have_state) return fail(c, RS_E_STATE, "state not set");
    return guarded(c, [&]() {
        HIPCHK(rs_launch_clamp(c->S[c->cur].p, (long long)c->N(), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return RS_OK;
    });
}

int rs_ssprk43_step(rs_ctx* c, double t, double dt, double atol, double rtol, double* eest) {
    (void)t;
    int rc = check_ready(c, false, true);
    if (rc) return rc;
    return guarded(c, [&]() {
        double* in = c->S[c->cur].p;
        double* outp = c->S[c->cur ^ 1].p;
        const bool keep = c->keep_stages && c->nt > 0 && trans_ppm(c) && ppm_vjp_sweeps(c);
        if (keep) {
            ppm_vjp_scratch(c);
            const PPMStages sg = ppm_ctx_stages(c);
            enqueue_ssprk43_ppm(c, nullptr, in, outp, dt, atol, rtol, 0, &sg);
        } else if (c->nt > 0) {
            enqueue_ssprk43(c, nullptr, in, outp, dt, atol, rtol, 0);
        }
        if (c->nt < c->ns) {
            const size_t off = (size_t)c->nt * c->nc;
            d2d(outp + off, in + off, (size_t)(c->ns - c->nt) * c->nc, c->stream);
        }
        double sse = c->nt > 0 ? host_sum(c, c->partial.p, (int)trans_nblocks(c)) : 0.0;
        if (eest) *eest = std::sqrt(sse / ((double)c->ns * (double)c->nc));
        c->have_cand = true;
        if (keep) {
            c->kept_ver = c->sver;
            c->kept_dt = dt;
        }
        return RS_OK;
    });
}

int rs_keep_transport_stages(rs_ctx* c, int keep) {
    if (!c) return RS_E_ARG;
    c->keep_stages = keep != 0;
    return RS_OK;
}

int rs_ros23_step(rs_ctx* c, double t, double dt, double atol, double rtol, double* eest) {
    (void)t;
    int rc = check_ready(c, true, false);
    if (rc) return rc;
    return guarded(c, [&]() {
        enqueue_ros23(c, nullptr, c->S[0].p, c->S[1].p, c->cur, dt, atol, rtol, 0);
        double sse = host_sum(c, c->partial.p, (int)chem_grid(c));
        if (eest) *eest = std::sqrt(sse / ((double)c->ns * (double)c->nc));
        c->have_cand = true;
        return RS_OK;
    });
}

int rs_ros23_step_noalloc(rs_ctx* c, const double* u_dev, double* unew_dev, long long ncells,
                          double dt, double atol, double rtol, double* eest) {
    if (!c) return RS_E_ARG;
    if (c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    if (!c->have_mech) return fail(c, RS_E_STATE, "mechanism not set");
    if (!u_dev || !unew_dev) return fail(c, RS_E_ARG, "null device buffer");
    if (ncells < 1 || ncells > c->nc)
        return fail(c, RS_E_ARG, "ncells must be in [1, NC] of the context's grid");
    const unsigned block = chem_threads(c);
    const long long grid = (ncells + block - 1) / block;
    if (grid > RS_NOALLOC_PARTIALS)
        return fail(c, RS_E_ARG, "ncells " + std::to_string(ncells) + " needs " +
                                     std::to_string(grid) + " error partials; the no-allocation "
                                     "entry point holds at most " +
                                     std::to_string(RS_NOALLOC_PARTIALS) + " (ncells <= " +
                                     std::to_string((long long)RS_NOALLOC_PARTIALS * block) +
                                     " at " + std::to_string(block) + " lanes per workgroup)");
    double* ptr = scratch_ptr(c->device);
    if (!ptr) return fail(c, RS_E_STATE, "no scratch on this device");
    return guarded(c, [&]() {
        RsChemArgs a{};
        a.S0 = const_cast<double*>(u_dev);   // read only: cur_static = 0 reads S0, writes S1
        a.S1 = unew_dev;
        a.ctrl = nullptr;
        a.K = c->K.p;
        a.jfac = c->has_jfac ? c->jfac.p : nullptr;
        a.E = c->has_E ? c->E.p : nullptr;
        a.partial = ptr;
        a.nc = ncells;
        a.ncol = c->ncol;
        a.dt_static = dt;
        a.atol = atol;
        a.rtol = rtol;
        // argument block on the stack: the struct, then the rate constants when
        // the kernel takes them inline (as chem_launch lays it out)
        const bool kin = !c->chem_kp && c->nr > 0 && c->nr <= RS_K_INLINE_MAX;
        alignas(8) char args[((sizeof(RsChemArgs) + 7) & ~(size_t)7) + RS_K_INLINE_MAX * sizeof(double)];
        const size_t off = (sizeof(RsChemArgs) + 7) & ~(size_t)7;
        size_t total = kin ? off + (size_t)c->nr * sizeof(double) : sizeof(RsChemArgs);
        std::memcpy(args, &a, sizeof(a));
        if (kin) std::memcpy(args + off, c->hK.data(), (size_t)c->nr * sizeof(double));
        void* cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, args, HIP_LAUNCH_PARAM_BUFFER_SIZE, &total,
                       HIP_LAUNCH_PARAM_END};
        HIPCHK(hipSetDevice(c->device));
        HIPCHK(hipModuleLaunchKernel(kin ? c->chem->attempt : c->chem->attempt_kp, (unsigned)grid,
                                     1, 1, block, 1, 1, (unsigned)chem_slab_bytes(c, block),
                                     c->stream, nullptr, cfg));
        // the sum of the partials in host_sum's order, folded inside the scratch
        const double* part = ptr;
        int n = (int)grid;
        if (n > 1024) {
            HIPCHK(rs_launch_sum(ptr, n, ptr + RS_NOALLOC_PARTIALS, c->stream));
            part = ptr + RS_NOALLOC_PARTIALS;
            n = rs_reduce_blocks(n);
        }
        HIPCHK(hipMemcpyAsync(c->hred, part, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double sse = 0.0;
        for (int i = 0; i < n; ++i) sse += c->hred[i];
        if (eest) *eest = std::sqrt(sse / ((double)c->ns * (double)ncells));
        return RS_OK;
    });
}

int rs_get_candidate(rs_ctx* c, double* u) {
    if (!c || !c->have_cand) return fail(c, RS_E_STATE, "no candidate");
    return guarded_keep(c, [&]() { d2h(u, c->S[c->cur ^ 1].p, c->N(), c->stream); return RS_OK; });
}

int rs_accept_candidate(rs_ctx* c, int clamp) {
    if (!c || !c->have_cand) return fail(c, RS_E_STATE, "no candidate");
    return guarded(c, [&]() {
        c->cur ^= 1;
        if (clamp_nonneg_of(clamp)) HIPCHK(rs_launch_clamp(c->S[c->cur].p, (long long)c->N(), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->have_cand = false;
        return RS_OK;
    });
}

int rs_adaptive_transport(rs_ctx* c, double t0, double t1, double* dt_io, const rs_pictrl* ctrl,
                          double atol, double rtol, int clamp, long long maxiters, rs_counts* out) {
    int rc = check_ready(c, false, true);
    if (rc) return rc;
    if (!dt_io || !ctrl || !(*dt_io > 0)) return fail(c, RS_E_ARG, "bad dt / ctrl");
    if (c->nt == 0) return fail(c, RS_E_STATE, "no transported species");
    c->have_cand = false;
    return guarded(c, [&]() {
        c->lastseq[0].clear();
        return run_adaptive(c, Op::T, t0, t1, dt_io, *ctrl, atol, rtol, clamp, maxiters, out,
                            &c->lastseq[0]);
    });
}

int rs_adaptive_chem(rs_ctx* c, double t0, double t1, double* dt_io, const rs_pictrl* ctrl,
                     double atol, double rtol, int clamp, long long maxiters, rs_counts* out) {
    int rc = check_ready(c, true, false);
    if (rc) return rc;
    if (!dt_io || !ctrl || !(*dt_io > 0)) return fail(c, RS_E_ARG, "bad dt / ctrl");
    c->have_cand = false;
    return guarded(c, [&]() {
        c->lastseq[1].clear();
        return run_adaptive(c, Op::C, t0, t1, dt_io, *ctrl, atol, rtol, clamp, maxiters, out,
                            &c->lastseq[1]);
    });
}

int rs_get_step_log(rs_ctx* c, int op, double* t_dt, long long cap, long long* n_out) {
    if (!c || (op != 0 && op != 1) || !n_out) return fail(c, RS_E_ARG, "bad step-log arguments");
    const auto& ts = c->lastseq[op];
    *n_out = (long long)ts.size();
    if (t_dt) {
        if (cap < (long long)ts.size()) return fail(c, RS_E_ARG, "step-log buffer too small");
        for (size_t i = 0; i < ts.size(); ++i) { t_dt[2 * i] = ts[i].first; t_dt[2 * i + 1] = ts[i].second; }
    }
    return RS_OK;
}

int rs_chem_percell(rs_ctx* c, double t0, double t1, double dt0, double* dt_cells,
                    const rs_pictrl* ctrl, double atol, double rtol, int clamp,
                    long long maxiters, rs_percell_stats* out) {
    int rc = check_ready(c, true, false);
    if (rc) return rc;
    if (!ctrl || !(dt0 > 0)) return fail(c, RS_E_ARG, "bad dt / ctrl");
    c->have_cand = false;
    return guarded(c, [&]() {
        return run_percell(c, t0, t1, dt0, dt_cells, *ctrl, atol, rtol, clamp, maxiters, out);
    });
}

int rs_chem_percell_vjp(rs_ctx* c, double t0, double t1, double dt0, const rs_pictrl* ctrl,
                        double atol, double rtol, int clamp, long long maxiters,
                        rs_percell_stats* out) {
    int rc = check_ready(c, true, false);
    if (rc) return rc;
    if (!ctrl || !(dt0 > 0)) return fail(c, RS_E_ARG, "bad dt / ctrl");
    c->have_cand = false;
    return guarded(c, [&]() {
        const size_t N = c->N();
        // a forward pass on a copy of the state finds the longest cell's
        // step count, which sizes the lane records
        c->pcu.ensure(N);
        d2d(c->pcu.p, c->S[c->cur].p, N, c->stream);
        rs_percell_stats fs{};
        const int r = run_percell(c, t0, t1, dt0, nullptr, *ctrl, atol, rtol, clamp, maxiters, &fs,
                                  c->pcu.p);
        if (r) return r;
        const long long nmax = fs.attempts_max + 1;
        ensure_percell_record(c, nmax);
        c->pcstats.ensure(9);
        enqueue_percell_adj(c, c->S[c->cur].p, nullptr, t0, t1, dt0, *ctrl, atol, rtol, clamp,
                            maxiters, nmax, c->pcstats.p);
        unsigned long long h[9];
        HIPCHK(hipMemcpyAsync(h, c->pcstats.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (out) {
            out->attempts_total = (long long)h[0];
            out->accepts_total = (long long)h[1];
            out->rejects_total = (long long)h[2];
            out->attempts_max = (long long)h[3];
            out->attempts_min = fs.attempts_min;
            out->n_failed = (long long)h[4];
        }
        if (h[6])
            return fail(c, RS_E_STATE, std::to_string(h[6]) + " cells took more steps than their "
                                       "forward pass (record overflow)");
        if (h[4]) return fail(c, RS_E_MAXITERS, std::to_string(h[4]) + " cells hit maxiters");
        return RS_OK;
    });
}

int rs_macro_step(rs_ctx* c, double t0, double t1, double* dtT, double* dtC,
                  const rs_split_opts* o, rs_counts* cT, rs_counts* cC) {
    int rc = check_ready(c, true, true);
    if (rc) return rc;
    if (!o || !dtT || !dtC) return fail(c, RS_E_ARG, "null argument");
    if (cT) *cT = rs_counts{};
    if (cC) *cC = rs_counts{};
    c->have_cand = false;
    return guarded(c, [&]() { return macro(c, t0, t1, dtT, dtC, *o, cT, cC, nullptr, nullptr); });
}

int rs_set_lambda(rs_ctx* c, const double* lam) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    return guarded_keep(c, [&]() {
        h2d(c->lam.p, lam, c->N(), c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        return RS_OK;
    });
}

int rs_get_lambda(rs_ctx* c, double* lam) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    return guarded_keep(c, [&]() { d2h(lam, c->lam.p, c->N(), c->stream); return RS_OK; });
}

int rs_zero_grad(rs_ctx* c) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    return guarded_keep(c, [&]() {
        HIPCHK(rs_launch_fill(c->gE.p, (long long)c->gE.n, 0.0, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return RS_OK;
    });
}

int rs_get_grad_emis(rs_ctx* c, double* g) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    return guarded_keep(c, [&]() { d2h(g, c->gE.p, c->gE.n, c->stream); return RS_OK; });
}

int rs_ssprk43_vjp(rs_ctx* c, double t, double dt) {
    (void)t;
    int rc = check_ready(c, false, true);
    if (rc) return rc;
    // PPM: the tendency is nonlinear, the VJP needs the stage states of the
    // step at the context's state; a step kept with rs_keep_transport_stages
    // at this state and dt left them in place (nothing has run since), and
    // a recompute leaves them for the next VJP of the same step
    const bool ppm = c->nt > 0 && trans_ppm(c);
    const bool reuse = ppm && c->kept_ver == c->sver && c->kept_dt == dt && ppm_vjp_sweeps(c);
    return guarded_keep(c, [&]() {
        if (ppm) {
            if (!reuse) ++c->sver;
            const PPMStages sg = ppm_ctx_stages(c);
            enqueue_ssprk43_vjp_ppm(c, c->S[c->cur].p, dt, reuse ? &sg : nullptr);
            if (ppm_vjp_sweeps(c)) {
                c->kept_ver = c->sver;
                c->kept_dt = dt;
            }
        } else if (c->nt > 0) {
            ++c->sver;
            enqueue_ssprk43_vjp(c, dt);
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        return RS_OK;
    });
}

int rs_ros23_vjp(rs_ctx* c, double t, double dt) {
    (void)t;
    int rc = check_ready(c, true, false);
    if (rc) return rc;
    return guarded_keep(c, [&]() {
        enqueue_ros23_vjp(c, c->S[c->cur].p, nullptr, dt);
        HIPCHK(hipStreamSynchronize(c->stream));
        return RS_OK;
    });
}

int rs_adjoint_run(rs_ctx* c, int nmacro, const double* tstops, double dtT0, double dtC0,
                   const rs_split_opts* o, int loss_kind, int loss_species, double* J_out,
                   double* grad_u0, double* grad_emis, rs_adjoint_stats* st) {
    int rc = check_ready(c, true, true);
    if (rc) return rc;
    if (!o || !tstops || nmacro < 1) return fail(c, RS_E_ARG, "bad adjoint arguments");
    rc = check_loss(c, loss_kind, loss_species);
    if (rc) return rc;
    c->have_cand = false;
    return guarded(c, [&]() -> int {
        using clk = std::chrono::steady_clock;
        rs_adjoint_stats S{};
        S.bad_op = S.bad_macro = -1;
        const size_t N = c->N();
        c->ckpt.ensure((size_t)(nmacro + 1) * N);
        c->seqT.assign(nmacro, {});
        c->seqC.assign(nmacro, {});
        c->rec_percell = o->chem_percell != 0;
        c->rec_dtC0 = dtC0;
        c->rec_pc_amax = 0;
        HIPCHK(hipStreamSynchronize(c->stream));
        auto tf0 = clk::now();
        double dtT = dtT0, dtC = dtC0;
        for (int m = 0; m < nmacro; ++m) {
            d2d(c->ckpt.p + (size_t)m * N, c->S[c->cur].p, N, c->stream);
            rs_counts cT{}, cC{};
            int r = macro(c, tstops[m], tstops[m + 1], &dtT, &dtC, *o, &cT, &cC, &c->seqT[m], &c->seqC[m]);
            if (r) return r;
            if (c->rec_percell) c->rec_pc_amax = std::max(c->rec_pc_amax, c->last_pc_amax);
            S.naccept_T += cT.naccept; S.nreject_T += cT.nreject;
            S.naccept_C += cC.naccept; S.nreject_C += cC.nreject;
        }
        d2d(c->ckpt.p + (size_t)nmacro * N, c->S[c->cur].p, N, c->stream);
        size_tape(c, nmacro);
        const int nvjp = sweep_vjps(c);
        c->nfc.ensure(2 * (size_t)(nvjp + 1));
        const double J = loss_on_device(c, c->S[c->cur].p, loss_kind, loss_species);
        if (J_out) *J_out = J;
        HIPCHK(hipStreamSynchronize(c->stream));
        S.seconds_forward = std::chrono::duration<double>(clk::now() - tf0).count();
        c->last_opts = *o;
        c->have_record = true;

        if (c->rec_percell) {
            ensure_percell_record(c, c->rec_pc_amax + 1);
            c->pcstats.ensure(9 * (size_t)nmacro);
        }
        auto tb0 = clk::now();
        backward_seed(c, loss_kind, loss_species, nvjp);
        c->relpart.ensure((size_t)nmacro * 1024);
        std::vector<hipEvent_t> ev(2 * (size_t)nmacro);
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        for (int m = nmacro - 1; m >= 0; --m)
            enqueue_backward_macro(c, m, *o, c->relpart.p + (size_t)m * 1024, ev[2 * m], ev[2 * m + 1]);
        HIPCHK(hipStreamSynchronize(c->stream));
        S.seconds_backward = std::chrono::duration<double>(clk::now() - tb0).count();
        double trep = 0.0;
        for (int m = 0; m < nmacro; ++m) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ev[2 * m], ev[2 * m + 1]));
            trep += ms * 1e-3;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        S.seconds_replay = trep;
        int worst = 0;
        S.replay_maxrel = host_max_many(c, c->relpart.p, nmacro, rs_reduce_blocks((long long)N),
                                        &worst);
        unsigned long long pc_ovf = 0, pc_fail = 0;
        if (c->rec_percell) {
            std::vector<unsigned long long> ps(9 * (size_t)nmacro);
            HIPCHK(hipMemcpyAsync(ps.data(), c->pcstats.p, ps.size() * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int m = 0; m < nmacro; ++m) {
                S.nvjp_C += (long long)ps[9 * (size_t)m + 5];
                pc_ovf += ps[9 * (size_t)m + 6];
                pc_fail += ps[9 * (size_t)m + 4];
            }
        }
        for (int m = 0; m < nmacro; ++m) {
            S.nvjp_T += (long long)c->seqT[m].size();
            if (!c->rec_percell) S.nvjp_C += (long long)c->seqC[m].size();
        }
        std::vector<unsigned long long> cnt(2 * (size_t)(nvjp + 1));
        HIPCHK(hipMemcpyAsync(cnt.data(), c->nfc.p, cnt.size() * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, c->stream));
        if (grad_u0) d2h(grad_u0, c->lam.p, N, c->stream);
        if (grad_emis) d2h(grad_emis, c->gE.p, c->gE.n, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        const bool bad = first_nonfinite(c, cnt, S);
        if (st) *st = S;
        if (pc_ovf || pc_fail)
            return fail(c, RS_E_STATE, "per-cell chemistry replay left its record (" +
                                       std::to_string(pc_ovf) + " cells past the recorded step "
                                       "count, " + std::to_string(pc_fail) + " at maxiters)");
        const std::string rf = replay_failure(S.replay_maxrel, worst);
        if (!rf.empty()) return fail(c, RS_E_STATE, rf);
        if (bad) return fail(c, RS_E_NONFINITE, nonfinite_message(S));
        return RS_OK;
    });
}

int rs_get_recorded_steps(rs_ctx* c, int macro_step, int op, double* t_dt, long long cap,
                          long long* n_out) {
    if (!c || !n_out || (op != 0 && op != 1)) return fail(c, RS_E_ARG, "bad arguments");
    if (!c->have_record || macro_step < 0 || macro_step >= (int)c->seqT.size())
        return fail(c, RS_E_STATE, "no recorded macro step " + std::to_string(macro_step));
    const auto& ts = (op ? c->seqC : c->seqT)[macro_step];
    if (op == 1 && c->rec_percell) {   // per-cell chemistry: no global step sequence
        *n_out = 0;
        return RS_OK;
    }
    *n_out = (long long)ts.size();
    if (t_dt) {
        if (cap < (long long)ts.size()) return fail(c, RS_E_ARG, "buffer too small");
        for (size_t i = 0; i < ts.size(); ++i) { t_dt[2 * i] = ts[i].first; t_dt[2 * i + 1] = ts[i].second; }
    }
    return RS_OK;
}

int rs_set_objective_weights(rs_ctx* c, const double* w) {
    if (!c || c->nc == 0) return fail(c, RS_E_STATE, "grid not set");
    return guarded(c, [&]() {
        if (!w) {
            c->have_wobj = false;
            return RS_OK;
        }
        c->wobj.resize(c->N());
        h2d(c->wobj.p, w, c->N(), c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        c->have_wobj = true;
        return RS_OK;
    });
}

int rs_replay_loss(rs_ctx* c, int loss_kind, int loss_species, double* J_out) {
    int rc = check_ready(c, true, true);
    if (rc) return rc;
    if (!c->have_record) return fail(c, RS_E_STATE, "no recorded sequence");
    rc = check_loss(c, loss_kind, loss_species);
    if (rc) return rc;
    return guarded(c, [&]() {
        const size_t N = c->N();
        const double* u = c->S[c->cur].p;
        // replay each macro step through the tape; the end state seeds the next
        DevArray<double> tmp;
        tmp.resize(N);
        d2d(tmp.p, u, N, c->stream);
        for (size_t m = 0; m < c->seqT.size(); ++m) {
            const size_t ns = replay_into_tape(c, tmp.p, c->seqT[m], c->seqC[m], c->last_opts);
            d2d(tmp.p, tape_ptr(c, ns * N) + (ns - 1) * N, N, c->stream);
        }
        const double J = loss_on_device(c, tmp.p, loss_kind, loss_species);
        if (J_out) *J_out = J;
        d2d(c->S[c->cur].p, tmp.p, N, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        return RS_OK;
    });
}

int rs_mechanism_info(rs_ctx* c, rs_mech_info* out) {
    if (!c || !c->have_mech || !out) return fail(c, RS_E_STATE, "mechanism not set");
    return guarded_keep(c, [&]() {
        const rs::MechPlan& p = c->chem->plan;
        rs_mech_info m{};
        m.ns = p.ns;
        m.nr = p.nr;
        m.nnz_J = (int)p.work.nnz_J;
        m.nnz_LU = (int)p.work.nnz_LU;
        m.ordering = p.ordering;
        m.flops_attempt = p.work.attempt();
        m.flops_vjp = p.work.vjp();
        hipFunction_t fs[3] = {c->chem->attempt, c->chem->percell, c->chem->vjp};
        int regs[3] = {0, 0, 0};
        long long scr[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i) {
            int v = 0;
            HIPCHK(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, fs[i]));
            regs[i] = v;
            HIPCHK(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, fs[i]));
            scr[i] = v;
        }
        m.regs_attempt = regs[0]; m.regs_percell = regs[1]; m.regs_vjp = regs[2];
        m.scratch_attempt = scr[0]; m.scratch_percell = scr[1]; m.scratch_vjp = scr[2];
        *out = m;
        return RS_OK;
    });
}

int rs_profile(rs_ctx* c, int enable) {
    if (!c) return RS_E_ARG;
    return guarded_keep(c, [&]() {
        if (enable && !c->prof) {
            prof_flush(c);
            for (int k = 0; k < P_COUNT; ++k) { c->prof_n[k] = 0; c->prof_ms[k] = 0; }
        }
        if (!enable && c->prof) prof_flush(c);
        c->prof = enable != 0;
        return RS_OK;
    });
}

int rs_profile_read(rs_ctx* c, rs_kernel_time* out, int max, int* n_out) {
    if (!c) return RS_E_ARG;
    return guarded_keep(c, [&]() {
        prof_flush(c);
        int n = 0;
        for (int k = 0; k < P_COUNT && n < max; ++k) {
            if (!c->prof_n[k]) continue;
            std::snprintf(out[n].name, sizeof(out[n].name), "%s", kProfNames[k]);
            out[n].launches = c->prof_n[k];
            out[n].total_ms = c->prof_ms[k];
            ++n;
        }
        if (n_out) *n_out = n;
        return RS_OK;
    });
}

}  // extern "C"

// ============================================================================
// Autotuner and ensembles
// ============================================================================
namespace {

double time_ms(rs_ctx* c, const std::function<void()>& body, int reps) {
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    body();   // warm-up
    HIPCHK(hipEventRecord(a, c->stream));
    for (int r = 0; r < reps; ++r) body();
    HIPCHK(hipEventRecord(b, c->stream));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / reps;
}

// A candidate replaces the incumbent only when at least 3% faster, so that
// timing noise between near-equal launch shapes does not flip the choice
// from run to run.
bool clearly_faster(double ms, double incumbent_ms) { return ms < 0.97 * incumbent_ms; }

int autotune(rs_ctx* c, rs_tune_result* out) {
    rs_tune_result r{};
    r.transport_tile_levels = c->tk;
    r.chem_block = (int)c->chem_block;
    const bool prof = c->prof;
    c->prof = false;
    double* in = c->S[c->cur].p;
    double* cand = c->S[c->cur ^ 1].p;
    if (c->nt > 0 && c->have_wind) {
        const int tks[4] = {4, 8, 16, 32};
        double best = 1e300;
        c->tmode = 0;
        for (int q = 0; q < 4; ++q) {
            c->tk = tks[q];
            r.transport_ms[q] = time_ms(c, [&]() {
                enqueue_ssprk43(c, nullptr, in, cand, 60.0, 1e-6, 1e-3, 0);
            }, 3);
            if (r.transport_ms[q] < best) { best = r.transport_ms[q]; r.transport_tile_levels = tks[q]; }
        }
        c->tk = r.transport_tile_levels;
        // the fused kernel with 1024- and 512-lane workgroups (two of the
        // smaller fit a CU, so one's barrier waits overlap the other's work).
        // The plan's own size is kept unless the other is clearly faster:
        // the two sizes sum the error norm over different blocks, so a
        // choice that flipped between runs on timing noise would change the
        // rounding of EEst, and with it a long run's step sequence.
        select_fused(c, 0);
        const int defnth = c->fnth;
        int bestf = 0;
        for (int nth : {defnth, defnth == 1024 ? 512 : 1024}) {
            select_fused(c, nth);
            if (!c->fplan_ok) continue;
            c->tmode = 1;
            // more repetitions than the per-stage candidates: the two sizes
            // are often within 10% of each other
            const double ms = time_ms(c, [&]() {
                enqueue_ssprk43(c, nullptr, in, cand, 60.0, 1e-6, 1e-3, 0);
            }, 10);
            if (bestf == 0 || clearly_faster(ms, r.transport_fused_ms)) {
                r.transport_fused_ms = ms;
                bestf = nth;
            }
        }
        const bool fused = bestf != 0 && r.transport_fused_ms < best;
        c->tmode = fused ? 1 : 0;
        c->fnth_req = fused ? bestf : 0;
        select_fused(c, c->fnth_req);
        r.transport_fused = trans_fused(c) ? 1 : 0;
        r.transport_fused_threads = fused ? bestf : 0;
    }
    if (c->have_mech) {
        // smallest block first; a larger one (another layout of the error
        // norm's partial sums) only when clearly faster, as above
        const unsigned bl[3] = {64, 128, 256};
        double best = 1e300;
        for (int q = 0; q < 3; ++q) {
            if (!chem_block_fits(c, bl[q])) continue;
            c->chem_block = bl[q];
            r.chem_ms[q] = time_ms(c, [&]() {
                enqueue_ros23(c, nullptr, c->S[0].p, c->S[1].p, c->cur, 1.0, 1e-6, 1e-3, 0);
            }, 3);
            if (best == 1e300 || clearly_faster(r.chem_ms[q], best)) {
                best = r.chem_ms[q];
                r.chem_block = (int)bl[q];
            }
        }
        c->chem_block = (unsigned)r.chem_block;
        // the same attempt with the rate constants through a pointer
        const bool kp0 = c->chem_kp;
        c->chem_kp = !kp0;
        const double alt = time_ms(c, [&]() {
            enqueue_ros23(c, nullptr, c->S[0].p, c->S[1].p, c->cur, 1.0, 1e-6, 1e-3, 0);
        }, 3);
        if (!clearly_faster(alt, best)) c->chem_kp = kp0;
        r.chem_kpointer = c->chem_kp ? 1 : 0;
        r.chem_kp_ms = kp0 ? best : alt;
    }
    c->prof = prof;
    c->have_cand = false;
    if (out) *out = r;
    return RS_OK;
}

}  // namespace

struct rs_ensemble {
    std::vector<rs_ctx*> m;
    DevArray<double> pool;       // replay tapes of one backward group, reused by every group
    int device = 0;
    hipStream_t origin = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    std::string err;
};

namespace {

int ens_fail(rs_ensemble* e, int code, const std::string& msg) {
    if (e) e->err = msg;
    return code;
}

// Captures body(d) for every member d in `idx` into one graph: the origin
// stream forks to each member's stream and joins them again.
hipGraphExec_t ens_capture(rs_ensemble* e, const std::vector<int>& idx,
                           const std::function<void(int)>& body, size_t* nnodes) {
    HIPCHK(hipStreamBeginCapture(e->origin, hipStreamCaptureModeGlobal));
    hipGraph_t g = nullptr;
    try {
        HIPCHK(hipEventRecord(e->fork, e->origin));
        for (int d : idx) {
            HIPCHK(hipStreamWaitEvent(e->m[d]->stream, e->fork, 0));
            body(d);
            HIPCHK(hipEventRecord(e->join[d], e->m[d]->stream));
            HIPCHK(hipStreamWaitEvent(e->origin, e->join[d], 0));
        }
    } catch (...) {
        (void)hipStreamEndCapture(e->origin, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    HIPCHK(hipStreamEndCapture(e->origin, &g));
    if (nnodes) {
        size_t n = 0;
        HIPCHK(hipGraphGetNodes(g, nullptr, &n));
        *nnodes = n;
    }
    hipGraphExec_t x = nullptr;
    HIPCHK(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    return x;
}

}  // namespace

extern "C" {

int rs_autotune(rs_ctx* c, rs_tune_result* out) {
    int rc = check_ready(c, false, false);
    if (rc) return rc;
    return guarded(c, [&]() { return autotune(c, out); });
}

int rs_ensemble_create(rs_ensemble** out, rs_ctx* const* members, int n) {
    if (!out || !members || n < 1) return RS_E_ARG;
    *out = nullptr;
    auto* e = new rs_ensemble();
    for (int d = 0; d < n; ++d) {
        if (!members[d]) { delete e; return RS_E_ARG; }
        if (d && members[d]->device != members[0]->device) { delete e; return RS_E_ARG; }
        e->m.push_back(members[d]);
    }
    e->device = members[0]->device;
    try {
        HIPCHK(hipSetDevice(e->device));
        HIPCHK(hipStreamCreateWithFlags(&e->origin, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
        e->join.resize(n);
        for (auto& j : e->join) HIPCHK(hipEventCreateWithFlags(&j, hipEventDisableTiming));
    } catch (const HipError&) {
        rs_ensemble_destroy(e);
        return RS_E_HIP;
    }
    *out = e;
    return RS_OK;
}

int rs_ensemble_destroy(rs_ensemble* e) {
    if (!e) return RS_OK;
    if (e->origin) (void)hipStreamSynchronize(e->origin);
    for (auto& j : e->join) if (j) (void)hipEventDestroy(j);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->origin) (void)hipStreamDestroy(e->origin);
    delete e;
    return RS_OK;
}

const char* rs_ensemble_last_error(const rs_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }

int rs_ensemble_autotune(rs_ensemble* e, rs_tune_result* out) {
    if (!e) return RS_E_ARG;
    for (size_t d = 0; d < e->m.size(); ++d) {
        int rc = rs_autotune(e->m[d], out ? out + d : nullptr);
        if (rc) return ens_fail(e, rc, "member " + std::to_string(d) + ": " + e->m[d]->err);
    }
    return RS_OK;
}

int rs_ensemble_adjoint_run(rs_ensemble* e, int nmacro, const double* tstops, double dtT0,
                            double dtC0, const rs_split_opts* o, int loss_kind, int loss_species,
                            double tape_budget_gib, double* J_out, rs_adjoint_stats* stats,
                            rs_ensemble_stats* es) {
    if (!e || !o || !tstops || nmacro < 1) return ens_fail(e, RS_E_ARG, "bad arguments");
    if (o->chem_percell) return ens_fail(e, RS_E_ARG, "adjoint needs the global-norm chemistry");
    const int n = (int)e->m.size();
    for (int d = 0; d < n; ++d) ++e->m[d]->sver;   // member states and scratch change
    for (int d = 0; d < n; ++d) {
        int rc = check_ready(e->m[d], true, true);
        if (!rc) rc = check_loss(e->m[d], loss_kind, loss_species);
        if (rc) return ens_fail(e, rc, "member " + std::to_string(d) + ": " + e->m[d]->err);
    }
    using clk = std::chrono::steady_clock;
    rs_ensemble_stats ES{};
    std::vector<rs_adjoint_stats> S(n);
    std::vector<int> profs(n);
    std::vector<int> nvjp(n, 0), worst(n, 0);
    // graph execs of the forward phases, destroyed on every path out
    struct ExecGuard {
        hipGraphExec_t x = nullptr;
        ~ExecGuard() { if (x) (void)hipGraphExecDestroy(x); }
    } gT, gC;
    try {
        HIPCHK(hipSetDevice(e->device));
        std::vector<int> all(n), withT;
        for (int d = 0; d < n; ++d) {
            rs_ctx* c = e->m[d];
            all[d] = d;
            if (c->nt > 0) withT.push_back(d);
            profs[d] = c->prof;
            c->prof = false;
            c->have_cand = false;
            const size_t N = c->N();
            c->ckpt.ensure((size_t)(nmacro + 1) * N);
            c->rec_percell = false;   // the ensemble runs the global-norm chemistry
            c->seqT.assign(nmacro, {});
            c->seqC.assign(nmacro, {});
            const long long cap = std::min<long long>(o->maxiters, 1 << 22);
            c->seqlog.ensure((size_t)(2 * std::max<long long>(cap, 1)));
            c->relpart.ensure((size_t)nmacro * 1024);
            if (trans_ppm(c) && c->nt > 0) ppm_vjp_scratch(c);
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        const int CH = 8;   // attempts per member per graph launch
        AdaptiveRun proto_T{Op::T, o->atol_T, o->rtol_T, o->clamp_nonneg & RS_CLAMP_NONNEG,
                            o->maxiters, 0, true};
        AdaptiveRun proto_C{Op::C, o->atol_C, o->rtol_C, o->clamp_nonneg, o->maxiters, 0, true};
        if (!withT.empty())
            gT.x = ens_capture(e, withT, [&](int d) { adaptive_chunk(e->m[d], proto_T, CH); }, nullptr);
        gC.x = ens_capture(e, all, [&](int d) { adaptive_chunk(e->m[d], proto_C, CH); }, nullptr);
        std::vector<double> dtT(n, dtT0), dtC(n, dtC0);
        auto tf0 = clk::now();
        auto phase = [&](Op op, hipGraphExec_t g, const std::vector<int>& idx, int mstep) {
            std::vector<AdaptiveRun> ar(n);
            for (int d : idx) {
                rs_ctx* c = e->m[d];
                const rs_pictrl& pc = (op == Op::T) ? o->ctrl_T : o->ctrl_C;
                ar[d] = adaptive_begin(c, op, tstops[mstep], tstops[mstep + 1],
                                       op == Op::T ? dtT[d] : dtC[d], pc,
                                       op == Op::T ? o->atol_T : o->atol_C,
                                       op == Op::T ? o->rtol_T : o->rtol_C,
                                       op == Op::T ? o->clamp_nonneg & RS_CLAMP_NONNEG
                                                   : o->clamp_nonneg,
                                       o->maxiters, true, e->origin);
            }
            for (;;) {
                bool done = true;
                for (int d : idx) done = done && e->m[d]->hctrl->done;
                if (done) break;
                HIPCHK(hipGraphLaunch(g, e->origin));
                ES.graph_launches_forward += 1;
                for (int d : idx)
                    HIPCHK(hipMemcpyAsync(e->m[d]->hctrl, e->m[d]->dctrl, sizeof(RsCtrl),
                                          hipMemcpyDeviceToHost, e->origin));
                HIPCHK(hipStreamSynchronize(e->origin));
            }
            for (int d : idx) {
                rs_ctx* c = e->m[d];
                rs_counts cnt{};
                auto* seq = (op == Op::T) ? &c->seqT[mstep] : &c->seqC[mstep];
                int rc = adaptive_finish(c, ar[d], op == Op::T ? &dtT[d] : &dtC[d], &cnt, seq);
                if (rc) throw ArgError{"member " + std::to_string(d) + ": " + c->err, rc};
                if (op == Op::T) { S[d].naccept_T += cnt.naccept; S[d].nreject_T += cnt.nreject; }
                else { S[d].naccept_C += cnt.naccept; S[d].nreject_C += cnt.nreject; }
            }
        };
        for (int mstep = 0; mstep < nmacro; ++mstep) {
            for (int d = 0; d < n; ++d) {
                rs_ctx* c = e->m[d];
                d2d(c->ckpt.p + (size_t)mstep * c->N(), c->S[c->cur].p, c->N(), c->stream);
                HIPCHK(hipStreamSynchronize(c->stream));
            }
            if (gT.x) phase(Op::T, gT.x, withT, mstep);
            phase(Op::C, gC.x, all, mstep);
        }
        for (int d = 0; d < n; ++d) {
            rs_ctx* c = e->m[d];
            d2d(c->ckpt.p + (size_t)nmacro * c->N(), c->S[c->cur].p, c->N(), c->stream);
            const double J = loss_on_device(c, c->S[c->cur].p, loss_kind, loss_species);
            if (J_out) J_out[d] = J;
            c->last_opts = *o;
            c->have_record = true;
            nvjp[d] = sweep_vjps(c);
            c->nfc.ensure(2 * (size_t)(nvjp[d] + 1));
        }
        ES.seconds_forward = std::chrono::duration<double>(clk::now() - tf0).count();
        if (gT.x) (void)hipGraphExecDestroy(gT.x);
        (void)hipGraphExecDestroy(gC.x);
        gT.x = gC.x = nullptr;

        // backward: groups of members whose tapes fit the budget
        auto tb0 = clk::now();
        for (int d = 0; d < n; ++d) {
            backward_seed(e->m[d], loss_kind, loss_species, nvjp[d]);
            HIPCHK(hipStreamSynchronize(e->m[d]->stream));
        }
        const double budget = (tape_budget_gib > 0 ? tape_budget_gib : 64.0) * 1073741824.0;
        std::vector<std::vector<int>> groups;
        double acc = 0.0;
        for (int d = 0; d < n; ++d) {
            const double b = (double)tape_bytes(e->m[d], nmacro);
            if (groups.empty() || (acc + b > budget && !groups.back().empty())) {
                groups.push_back({});
                acc = 0.0;
            }
            groups.back().push_back(d);
            acc += b;
        }
        // one pool, sized for the largest group, lent to the members of each
        // group in turn (no allocation between groups)
        size_t pool_n = 0;
        for (auto& grp : groups) {
            size_t g = 0;
            for (int d : grp) g += tape_bytes(e->m[d], nmacro) / sizeof(double);
            pool_n = std::max(pool_n, g);
        }
        auto ts0 = clk::now();
        e->pool.ensure(pool_n);
        ES.seconds_backward_setup += std::chrono::duration<double>(clk::now() - ts0).count();
        for (auto& grp : groups) {
            auto tg0 = clk::now();
            size_t off = 0;
            for (int d : grp) {
                rs_ctx* c = e->m[d];
                const size_t need = tape_bytes(c, nmacro) / sizeof(double);
                c->tape_view = e->pool.p + off;
                c->tape_view_n = need;
                off += need;
            }
            size_t nn = 0;
            ExecGuard gB;
            gB.x = ens_capture(e, grp, [&](int d) {
                rs_ctx* c = e->m[d];
                for (int mstep = nmacro - 1; mstep >= 0; --mstep)
                    enqueue_backward_macro(c, mstep, *o, c->relpart.p + (size_t)mstep * 1024, nullptr, nullptr);
            }, &nn);
            auto tg1 = clk::now();
            ES.seconds_backward_setup += std::chrono::duration<double>(tg1 - tg0).count();
            HIPCHK(hipGraphLaunch(gB.x, e->origin));
            HIPCHK(hipStreamSynchronize(e->origin));
            ES.seconds_backward_run += std::chrono::duration<double>(clk::now() - tg1).count();
            ES.graph_launches_backward += 1;
            ES.graph_nodes_backward += (long long)nn;
            for (int d : grp) {
                rs_ctx* c = e->m[d];
                S[d].replay_maxrel = host_max_many(c, c->relpart.p, nmacro,
                                                   rs_reduce_blocks((long long)c->N()), &worst[d]);
                for (int mstep = 0; mstep < nmacro; ++mstep) {
                    S[d].nvjp_T += (long long)c->seqT[mstep].size();
                    S[d].nvjp_C += (long long)c->seqC[mstep].size();
                }
                c->tape_view = nullptr;
                c->tape_view_n = 0;
            }
        }
        ES.backward_groups = (int)groups.size();
        ES.seconds_backward = std::chrono::duration<double>(clk::now() - tb0).count();
    } catch (const HipError& x) {
        for (int d = 0; d < n; ++d) { e->m[d]->prof = profs[d]; e->m[d]->tape_view = nullptr; }
        return ens_fail(e, RS_E_HIP, std::string(x.what) + ": " + hipGetErrorString(x.e));
    } catch (const ArgError& x) {
        for (int d = 0; d < n; ++d) { e->m[d]->prof = profs[d]; e->m[d]->tape_view = nullptr; }
        return ens_fail(e, x.code, x.msg);
    }
    int rc = RS_OK;
    std::string msg;
    for (int d = 0; d < n; ++d) {
        rs_ctx* c = e->m[d];
        c->prof = profs[d];
        S[d].seconds_forward = ES.seconds_forward;
        S[d].seconds_backward = ES.seconds_backward;
        bool bad = false;
        try {
            std::vector<unsigned long long> cnt(2 * (size_t)(nvjp[d] + 1));
            HIPCHK(hipMemcpy(cnt.data(), c->nfc.p, cnt.size() * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost));
            bad = first_nonfinite(c, cnt, S[d]);
        } catch (const HipError& x) {
            return ens_fail(e, RS_E_HIP, std::string(x.what) + ": " + hipGetErrorString(x.e));
        }
        if (stats) stats[d] = S[d];
        if (rc) continue;
        const std::string rf = replay_failure(S[d].replay_maxrel, worst[d]);
        if (!rf.empty()) {
            rc = RS_E_STATE;
            msg = "member " + std::to_string(d) + ": " + rf;
        } else if (bad) {
            rc = RS_E_NONFINITE;
            msg = "member " + std::to_string(d) + ": " + nonfinite_message(S[d]);
        }
    }
    if (es) *es = ES;
    if (rc) return ens_fail(e, rc, msg);
    return RS_OK;
}

}  // extern "C"
