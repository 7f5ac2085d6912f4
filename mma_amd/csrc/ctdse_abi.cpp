// Host side of libsingleTDSE.so: the reference's C entry points and the batched
// extension, orchestrating the HIP kernels of tdse_setup.hip / tdse_propagate.hip.
// Every compute step runs on the GPU; the host only validates arguments, moves
// data across the boundary and owns the reference's malloc/calloc output arrays.
#include "ctdse.h"
#include "tdse_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

// Device memory of ctdse_propagate_one_noalloc, one copy per GPU for the life of the process.
// Host code never uses this name as an address: the device address comes from
// hipGetSymbolAddress (scratch_address below).
static __device__ __align__(256) unsigned char scratch[CTDSE_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local int g_status = CTDSE_OK;
thread_local std::string g_message;

int set_status(int code, const std::string &msg)
{
    g_status = code;
    g_message = msg;
    return code;
}
int clear_status()
{
    g_status = CTDSE_OK;
    g_message.clear();
    return CTDSE_OK;
}

#define HIP_CHECK(expr)                                                                                   \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return set_status(CTDSE_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device allocation owned by the holder.
struct DevMem {
    void *p = nullptr;
    DevMem() = default;
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    ~DevMem() { reset(); }
    void reset()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    hipError_t alloc(size_t bytes)
    {
        reset();
        return hipMalloc(&p, bytes ? bytes : 8);
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

int device_ready()
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return set_status(CTDSE_ERR_DEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
    return CTDSE_OK;
}

// Persistent-kernel workspace (one psi buffer per workgroup slot, updated in place), grown on demand.
std::mutex g_launch_mutex;
DevMem g_work;
size_t g_work_bytes = 0;

constexpr int kGsMaxIter = 1000;      // tridiag.c:139
constexpr int kDefaultStepsPerPass = 1;  // propagation kernel used when no launch shape is given
constexpr double kTruncTol = 1e-17;   // relative truncation of the segmented solve

} // namespace

// Lattice tables shared by geometries of one batch (TdseGeom::sg): g, V' and 1/beta of the
// widest geometry in eight alignments (shift 0..7 points), so that every geometry's chunks
// start on a shared chunk boundary.
struct Lattice {
    static constexpr int kAlign = TDSE_CHUNK;
    DevMem g[kAlign], dpot[kAlign], ib[kAlign], chq[kAlign], chg[kAlign];
};

struct ctdse_geometry {
    TdseGeom dev{};
    int num_r = 0;
    double dx = 0, dt = 0, a = 0, x_int = 0;
    double e_init = 0;
    int gs_iters = 0;
    DevMem x, psi0_nat, g, dpot, ib, psi0, chq, chg, adq, adg;
    std::shared_ptr<Lattice> lattice;  // null, or the batch's shared tables this geometry reads
    int lattice_shift = -1;            // the alignment of `lattice` it reads (0..7)
};

namespace {

constexpr size_t kSetupScratchBytes = size_t(8) << 30;  // device scratch per group of setup jobs
constexpr int kSetupGroupJobs = 16384;                      // and at most this many jobs per group

// Runs ground-state jobs (J.x and J.psi0 set) in groups of similar size, largest first: a
// group's launch lasts as long as its largest serial inverse iteration.
int run_gs_jobs(std::vector<GsJob> &jobs)
{
    std::vector<size_t> idx(jobs.size());
    for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t p, size_t q) { return jobs[p].num_r > jobs[q].num_r; });
    size_t i0 = 0;
    while (i0 < idx.size()) {
        size_t i1 = i0, bytes = 0;
        while (i1 < idx.size() && (i1 - i0) < (size_t)kSetupGroupJobs) {
            const size_t b = 16 * ((size_t)jobs[idx[i1]].num_r + 2) * sizeof(double);
            if (i1 > i0 && bytes + b > kSetupScratchBytes) break;
            bytes += b;
            i1++;
        }
        const int m = (int)(i1 - i0);
        DevMem work, d_jobs;
        HIP_CHECK(work.alloc(bytes));
        HIP_CHECK(d_jobs.alloc(m * sizeof(GsJob)));
        std::vector<GsJob> group(m);
        size_t off = 0;
        for (int k = 0; k < m; k++) {
            group[k] = jobs[idx[i0 + k]];
            group[k].max_iter = kGsMaxIter;
            group[k].work = work.as<double>() + off;
            off += 16 * ((size_t)group[k].num_r + 2);
        }
        HIP_CHECK(hipMemcpy(d_jobs.p, group.data(), m * sizeof(GsJob), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(tdse_ground_state_kernel, dim3(m), dim3(64), 0, 0, d_jobs.as<GsJob>(), m);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpy(group.data(), d_jobs.p, m * sizeof(GsJob), hipMemcpyDeviceToHost));
        for (int k = 0; k < m; k++) jobs[idx[i0 + k]] = group[k];
        i0 = i1;
    }
    return CTDSE_OK;
}

int gs_status(const GsJob &J, const std::string &who)
{
    if (J.status == 1)
        return set_status(CTDSE_ERR_GS_NOT_CONVERGED, who + "ground state: |dE| > CV after 1000 inverse iterations");
    if (J.status == 2) return set_status(CTDSE_ERR_ZERO_PIVOT, who + "ground state: zero pivot in the Thomas solve");
    return CTDSE_OK;
}

TableJob table_job(ctdse_geometry *G)
{
    TableJob J{};
    J.num_r = G->num_r;
    J.dx = G->dx;
    J.dt = G->dt;
    J.a = G->a;
    J.x_int = G->x_int;
    J.x = G->x.as<double>();
    J.psi0_nat = G->psi0_nat.as<double>();
    J.g = G->g.as<double>();
    J.dpot = G->dpot.as<double>();
    J.ib = G->ib.as<double2>();
    J.psi0 = G->psi0.as<double2>();
    J.chq = G->chq.as<double2>();
    J.chg = G->chg.as<double2>();
    J.adq = G->adq.as<double2>();
    J.adg = G->adg.as<double2>();
    return J;
}

int alloc_tables(ctdse_geometry *G)
{
    const int nx = G->num_r + 1;
    const int nchunks = (nx + TDSE_CHUNK - 1) / TDSE_CHUNK;
    const long long npad = tdse_tiled_size(nchunks);
    HIP_CHECK(G->g.alloc(npad * sizeof(double)));
    HIP_CHECK(G->dpot.alloc(npad * sizeof(double)));
    HIP_CHECK(G->ib.alloc(npad * sizeof(double2)));
    HIP_CHECK(G->psi0.alloc(npad * sizeof(double2)));
    HIP_CHECK(G->chq.alloc((nchunks + 128) * sizeof(double2)));
    HIP_CHECK(G->chg.alloc((nchunks + 128) * sizeof(double2)));
    HIP_CHECK(G->adq.alloc((nchunks + 128) * sizeof(double2)));
    HIP_CHECK(G->adg.alloc((nchunks + 128) * sizeof(double2)));
    return CTDSE_OK;
}

// Host part of the table build once the kernels have run: truncation depth of the
// segmented solve and the device descriptor.
int finish_tables(ctdse_geometry *G, const TableJob &J, const std::string &who)
{
    const int nx = G->num_r + 1;
    const int nchunks = (nx + TDSE_CHUNK - 1) / TDSE_CHUNK;
    const double rho = std::max(J.rho_fwd, J.rho_bwd);
    if (!(rho > 0.0) || !(rho < 0.999) || !std::isfinite(rho))
        return set_status(CTDSE_ERR_UNSUPPORTED, who + "Crank-Nicolson factor decay rate " + std::to_string(rho) +
                                                     " too close to 1 for the segmented solve");
    const int W = (int)std::ceil(std::log(kTruncTol) / std::log(rho));
    const int H = (W + TDSE_CHUNK - 1) / TDSE_CHUNK;
    if (H > TDSE_MAX_HALO || TDSE_MAX_WAVES * 64 - H < 64)
        return set_status(CTDSE_ERR_UNSUPPORTED, who + "truncation depth " + std::to_string(W) +
                                                     " points exceeds the segment size (dt/dx^2 too large)");
    if (J.pint_lo < 0 || J.pint_hi > G->num_r) return set_status(CTDSE_ERR_ARGUMENT, who + "x_int lies outside the grid");
    TdseGeom &D = G->dev;
    D.g = G->g.as<double>();
    D.ib = G->ib.as<double2>();
    D.dpot = G->dpot.as<double>();
    D.psi0 = G->psi0.as<double2>();
    D.chq = G->chq.as<double2>();
    D.chg = G->chg.as<double2>();
    D.adq = G->adq.as<double2>();
    D.adg = G->adg.as<double2>();
    D.nx = nx;
    D.nchunks = nchunks;
    D.npad = tdse_tiled_size(nchunks);
    D.halo_lanes = H;
    D.pint_lo = J.pint_lo;
    D.pint_hi = J.pint_hi;
    D.p0_lo = J.p0_lo;
    D.p0_hi = J.p0_hi;
    D.dx = G->dx;
    D.dt = G->dt;
    D.xmax = 0.5 * G->num_r * G->dx;
    D.k36 = 36.0 * G->dt / (G->dx * G->dx);
    D.pint0 = J.pint0;
    return CTDSE_OK;
}

// Host copies of one geometry's grid and ground state (natural order, interleaved).
struct HostGrid {
    std::vector<double> x, psi0;
};

// Builds the factor tables of several geometries (x and psi0_nat set) in grouped launches;
// note(index[k], code) is called (with the message set) for each geometry that fails.
// host: one geometry alone with its grid and ground state on the host; its serial pivot scan
// then runs on a host core (a few ms for 16001 points, against ~24 ms in one GPU lane), with
// the same bits.
template <class Note>
int build_tables(const std::vector<ctdse_geometry *> &gs, const std::vector<int> &index, Note note,
                 const HostGrid *host = nullptr)
{
    if (host && gs.size() == 1) {
        ctdse_geometry *G = gs[0];
        const size_t nx = G->num_r + 1;
        DevMem piv, d_job;
        if (alloc_tables(G) != CTDSE_OK) return g_status;
        HIP_CHECK(piv.alloc(2 * nx * sizeof(double)));
        HIP_CHECK(d_job.alloc(sizeof(TableJob)));
        std::vector<double> piv_h(2 * nx);
        TableJob Jh = table_job(G);
        Jh.x = host->x.data();
        Jh.psi0_nat = host->psi0.data();
        Jh.pivot_work = piv_h.data();
        tdse_tables_pivot_host(Jh);
        TableJob J = table_job(G);
        J.pivot_work = piv.as<double>();
        J.rho_fwd = Jh.rho_fwd;
        J.rho_bwd = Jh.rho_bwd;
        J.pint_lo = Jh.pint_lo;
        J.pint_hi = Jh.pint_hi;
        J.p0_lo = Jh.p0_lo;
        J.p0_hi = Jh.p0_hi;
        J.pint0 = Jh.pint0;
        HIP_CHECK(hipMemcpy(piv.p, piv_h.data(), 2 * nx * sizeof(double), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(d_job.p, &J, sizeof(TableJob), hipMemcpyHostToDevice));
        const int nchunks = (int)((nx + TDSE_CHUNK - 1) / TDSE_CHUNK);
        hipLaunchKernelGGL(tdse_tables_grid_kernel, dim3(64, 1), dim3(256), 0, 0, d_job.as<TableJob>());
        hipLaunchKernelGGL(tdse_tables_ib_kernel, dim3((unsigned)((nx + 255) / 256), 1), dim3(256), 0, 0,
                           d_job.as<TableJob>());
        hipLaunchKernelGGL(tdse_tables_chunk_kernel, dim3((nchunks + 255) / 256, 1), dim3(256), 0, 0,
                           d_job.as<TableJob>());
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(nullptr));  // the scratch buffers are freed on return
        const int st = finish_tables(G, J, "");
        if (st != CTDSE_OK) note(index[0], st);
        return CTDSE_OK;
    }
    size_t i0 = 0;
    while (i0 < gs.size()) {
        size_t i1 = i0, bytes = 0;
        int max_chunks = 1;
        while (i1 < gs.size() && (i1 - i0) < (size_t)kSetupGroupJobs) {
            const size_t b = 2 * ((size_t)gs[i1]->num_r + 1) * sizeof(double);
            if (i1 > i0 && bytes + b > kSetupScratchBytes) break;
            bytes += b;
            max_chunks = std::max(max_chunks, (gs[i1]->num_r + TDSE_CHUNK) / TDSE_CHUNK);
            i1++;
        }
        const int m = (int)(i1 - i0);
        DevMem piv, d_jobs;
        HIP_CHECK(piv.alloc(bytes));
        HIP_CHECK(d_jobs.alloc(m * sizeof(TableJob)));
        std::vector<TableJob> jobs(m);
        size_t off = 0;
        for (int k = 0; k < m; k++) {
            ctdse_geometry *G = gs[i0 + k];
            if (alloc_tables(G) != CTDSE_OK) return g_status;
            jobs[k] = table_job(G);
            jobs[k].pivot_work = piv.as<double>() + off;
            off += 2 * ((size_t)G->num_r + 1);
        }
        HIP_CHECK(hipMemcpy(d_jobs.p, jobs.data(), m * sizeof(TableJob), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(tdse_tables_grid_kernel, dim3(64, m), dim3(256), 0, 0, d_jobs.as<TableJob>());
        hipLaunchKernelGGL(tdse_tables_pivot_kernel, dim3(m), dim3(64), 0, 0, d_jobs.as<TableJob>());
        hipLaunchKernelGGL(tdse_tables_ib_kernel, dim3((unsigned)((max_chunks * TDSE_CHUNK + 255) / 256), m), dim3(256),
                           0, 0, d_jobs.as<TableJob>());
        hipLaunchKernelGGL(tdse_tables_chunk_kernel, dim3((max_chunks + 255) / 256, m), dim3(256), 0, 0,
                           d_jobs.as<TableJob>());
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpy(jobs.data(), d_jobs.p, m * sizeof(TableJob), hipMemcpyDeviceToHost));
        for (int k = 0; k < m; k++) {
            const int i = index[i0 + k];
            const std::string who = gs.size() > 1 ? "geometry " + std::to_string(i) + ": " : "";
            const int st = finish_tables(gs[i0 + k], jobs[k], who);
            if (st != CTDSE_OK) note(i, st);
        }
        i0 = i1;
    }
    return CTDSE_OK;
}

struct GeomSpec {
    int num_r;
    double dx, dt, a, e_guess, cv, x_int;
    const double *psi0_host;  // null: ground state by inverse iteration
    const double *x_host;     // null: x_k = k dx - num_r dx / 2
};

int check_spec(const GeomSpec &g, const std::string &who)
{
    if (g.num_r < 2) return set_status(CTDSE_ERR_ARGUMENT, who + "num_r must be >= 2");
    if (!(g.dx > 0) || !(g.dt > 0) || !(g.a > 0) || !std::isfinite(g.dx) || !std::isfinite(g.dt) || !std::isfinite(g.a))
        return set_status(CTDSE_ERR_ARGUMENT, who + "dx, dt and the soft-Coulomb parameter a must be positive");
    if (!(g.x_int > 0) || !std::isfinite(g.x_int)) return set_status(CTDSE_ERR_ARGUMENT, who + "x_int must be positive");
    return CTDSE_OK;
}

// Grid and ground state of one geometry on the host: x_k = k dx - num_r dx / 2 (tools.c:62-79)
// unless given, psi0 given or by the inverse iteration (tdse_ground_state_host, the same bits
// as the reference).  Sets G's energy and iteration count; CTDSE_ERR_GS_NOT_CONVERGED still
// leaves the last iterate in hg.
int host_grid(const GeomSpec &g, ctdse_geometry *G, HostGrid &hg)
{
    const size_t nx = g.num_r + 1;
    hg.x.resize(nx);
    hg.psi0.resize(2 * nx);
    if (g.x_host) {
        std::copy(g.x_host, g.x_host + nx, hg.x.begin());
    } else {
        const double half_span = 0.5 * g.num_r * g.dx;
        for (size_t k = 0; k < nx; k++) hg.x[k] = (double)k * g.dx - half_span;
    }
    if (g.psi0_host) {
        std::copy(g.psi0_host, g.psi0_host + 2 * nx, hg.psi0.begin());
        return CTDSE_OK;
    }
    std::vector<double> work(16 * ((size_t)g.num_r + 2));
    GsJob J{};
    J.num_r = g.num_r;
    J.max_iter = kGsMaxIter;
    J.dx = g.dx;
    J.a = g.a;
    J.e_guess = g.e_guess;
    J.cv = g.cv;
    J.x = hg.x.data();  // the iteration writes the same grid
    J.psi0 = hg.psi0.data();
    J.work = work.data();
    tdse_ground_state_host(J);
    G->e_init = J.energy;
    G->gs_iters = J.iters;
    return gs_status(J, "");
}

int upload_grid(ctdse_geometry *G, const HostGrid &hg, const std::string &who)
{
    hipError_t e = G->x.alloc(hg.x.size() * sizeof(double));
    if (e == hipSuccess) e = G->psi0_nat.alloc(hg.psi0.size() * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(G->x.p, hg.x.data(), hg.x.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(G->psi0_nat.p, hg.psi0.data(), hg.psi0.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return set_status(CTDSE_ERR_DEVICE, who + "geometry upload: " + hipGetErrorString(e));
    return CTDSE_OK;
}

// Geometries of one batch that lie on one lattice (equal dx, dt and a, and num_r of one
// parity, so x_j = (j - num_r/2) dx up to rounding) read g, V' and 1/beta from one shared set
// of tables, the widest geometry's: a batch of many grid sizes then streams one geometry's
// tables instead of one per member.  Geometry i of the group starts off = (N - n_i)/2 points
// into the widest one (N points); it reads the alignment shifted by (-off) mod 8 points, at a
// chunk offset of (off + shift)/8.  Its first H chunks keep their own pivots (the pivots of
// a grid that starts later differ from the widest grid's near that start; the difference
// decays by ~rho^2 per point, below 1e-30 after H chunks).
int share_lattices(const std::vector<ctdse_geometry *> &gs)
{
    std::map<std::tuple<double, double, double, int>, std::vector<ctdse_geometry *>> groups;
    for (ctdse_geometry *G : gs)
        if (G && G->dev.g) groups[std::make_tuple(G->dx, G->dt, G->a, G->num_r & 1)].push_back(G);
    for (auto &kv : groups) {
        const std::vector<ctdse_geometry *> &grp = kv.second;
        if (grp.size() < 2) continue;
        const ctdse_geometry *big = *std::max_element(
            grp.begin(), grp.end(), [](const ctdse_geometry *p, const ctdse_geometry *q) { return p->num_r < q->num_r; });
        auto L = std::make_shared<Lattice>();
        const long long npad = tdse_tiled_size(big->dev.nchunks + 64);
        for (int s = 0; s < Lattice::kAlign; s++) {
            HIP_CHECK(L->g[s].alloc(npad * sizeof(double)));
            HIP_CHECK(L->dpot[s].alloc(npad * sizeof(double)));
            HIP_CHECK(L->ib[s].alloc(npad * sizeof(double2)));
            hipLaunchKernelGGL(tdse_lattice_kernel, dim3((unsigned)((npad + 255) / 256)), dim3(256), 0, 0,
                               big->g.as<double>(), big->dpot.as<double>(), big->ib.as<double2>(), big->dev.nx, s,
                               npad, L->g[s].as<double>(), L->dpot[s].as<double>(), L->ib[s].as<double2>());
            // chunk maps of the shared chunks that hold lattice rows (chunk k at [k + 64])
            const int nch = big->dev.nchunks + 1;
            HIP_CHECK(L->chq[s].alloc((nch + 128) * sizeof(double2)));
            HIP_CHECK(L->chg[s].alloc((nch + 128) * sizeof(double2)));
            hipLaunchKernelGGL(tdse_lattice_chunk_kernel, dim3((unsigned)((nch + 255) / 256)), dim3(256), 0, 0,
                               L->g[s].as<double>(), L->ib[s].as<double2>(), nch, L->chq[s].as<double2>(),
                               L->chg[s].as<double2>());
        }
        HIP_CHECK(hipGetLastError());
        for (ctdse_geometry *G : grp) {
            const int off = (big->num_r - G->num_r) / 2;
            const int shift = (Lattice::kAlign - off % Lattice::kAlign) % Lattice::kAlign;
            G->lattice = L;
            G->lattice_shift = shift;
            G->dev.sg = L->g[shift].as<double>();
            G->dev.sdp = L->dpot[shift].as<double>();
            G->dev.sib = L->ib[shift].as<double2>();
            G->dev.schq = L->chq[shift].as<double2>();
            G->dev.schg = L->chg[shift].as<double2>();
            G->dev.s_off = (off + shift) / TDSE_CHUNK;
            G->dev.head_chunks = G->dev.halo_lanes;
        }
    }
    HIP_CHECK(hipStreamSynchronize(nullptr));
    return CTDSE_OK;
}

// Builds geometries: ground states of all that need one in grouped launches, then the
// tables.  out[i] is null where status[i] is a hard failure; a ground state that did not
// converge (CTDSE_ERR_GS_NOT_CONVERGED) still yields a geometry, as in the reference, which
// carries on with the last iterate.  Returns the first non-OK status (or OK).
int geometries_from(const std::vector<GeomSpec> &spec, std::vector<ctdse_geometry *> &out, std::vector<int> &status)
{
    const int n = (int)spec.size();
    out.assign(n, nullptr);
    status.assign(n, CTDSE_OK);
    if (n == 0) return CTDSE_OK;
    if (device_ready() != CTDSE_OK) {
        status.assign(n, g_status);
        return g_status;
    }
    std::string first_msg;
    int first = CTDSE_OK;
    auto note = [&](int i, int code) {
        status[i] = code;
        if (first == CTDSE_OK && code != CTDSE_OK) {
            first = code;
            first_msg = g_message;
        }
    };
    auto who = [&](int i) { return n > 1 ? "geometry " + std::to_string(i) + ": " : std::string(); };
    std::vector<GsJob> jobs;
    std::vector<int> job_of;
    // one geometry alone: its serial steps (ground state, pivots) run on the host, where one
    // core is far faster than one GPU lane; many geometries run them side by side on the GPU
    const bool on_host = n == 1;
    HostGrid hg;
    for (int i = 0; i < n; i++) {
        const GeomSpec &g = spec[i];
        if (check_spec(g, who(i)) != CTDSE_OK) {
            note(i, g_status);
            continue;
        }
        auto *G = new ctdse_geometry;
        out[i] = G;
        G->num_r = g.num_r;
        G->dx = g.dx;
        G->dt = g.dt;
        G->a = g.a;
        G->x_int = g.x_int;
        const size_t nx = g.num_r + 1;
        if (on_host) {
            const int gst = host_grid(g, G, hg);
            if (gst != CTDSE_OK) note(i, gst);
            if (gst != CTDSE_OK && gst != CTDSE_ERR_GS_NOT_CONVERGED) continue;
            const int st = upload_grid(G, hg, who(i));
            if (st != CTDSE_OK) note(i, st);
            continue;
        }
        hipError_t e = G->x.alloc(nx * sizeof(double));
        if (e == hipSuccess) e = G->psi0_nat.alloc(2 * nx * sizeof(double));
        if (e == hipSuccess && g.psi0_host) {
            e = hipMemcpy(G->psi0_nat.p, g.psi0_host, 2 * nx * sizeof(double), hipMemcpyHostToDevice);
            if (e == hipSuccess && g.x_host) e = hipMemcpy(G->x.p, g.x_host, nx * sizeof(double), hipMemcpyHostToDevice);
        }
        if (e == hipSuccess && (!g.psi0_host || !g.x_host)) {
            // the ground-state kernel writes x itself; a given psi0 without x gets the grid here
            if (g.psi0_host) {
                hipLaunchKernelGGL(tdse_xgrid_kernel, dim3((nx + 255) / 256), dim3(256), 0, 0, G->x.as<double>(),
                                   g.num_r, g.dx);
                e = hipGetLastError();
            }
        }
        if (e != hipSuccess) {
            set_status(CTDSE_ERR_DEVICE, who(i) + "geometry upload: " + hipGetErrorString(e));
            note(i, g_status);
            continue;
        }
        if (!g.psi0_host) {
            GsJob J{};
            J.num_r = g.num_r;
            J.dx = g.dx;
            J.a = g.a;
            J.e_guess = g.e_guess;
            J.cv = g.cv;
            J.x = G->x.as<double>();
            J.psi0 = G->psi0_nat.as<double>();
            jobs.push_back(J);
            job_of.push_back(i);
        }
    }
    if (!jobs.empty()) {
        if (run_gs_jobs(jobs) != CTDSE_OK) {
            for (int i = 0; i < n; i++) {
                delete out[i];
                out[i] = nullptr;
                status[i] = g_status;
            }
            return g_status;
        }
        for (size_t k = 0; k < jobs.size(); k++) {
            ctdse_geometry *G = out[job_of[k]];
            G->e_init = jobs[k].energy;
            G->gs_iters = jobs[k].iters;
            const int st = gs_status(jobs[k], who(job_of[k]));
            if (st != CTDSE_OK) note(job_of[k], st);
        }
    }
    std::vector<ctdse_geometry *> tab;
    std::vector<int> tab_index;
    for (int i = 0; i < n; i++)
        if (out[i] && (status[i] == CTDSE_OK || status[i] == CTDSE_ERR_GS_NOT_CONVERGED)) {
            tab.push_back(out[i]);
            tab_index.push_back(i);
        }
    if (build_tables(tab, tab_index, note, on_host ? &hg : nullptr) != CTDSE_OK) {
        for (int i = 0; i < n; i++) {
            delete out[i];
            out[i] = nullptr;
            status[i] = g_status;
        }
        return g_status;
    }
    for (int i = 0; i < n; i++) {
        if (out[i] && status[i] != CTDSE_OK && status[i] != CTDSE_ERR_GS_NOT_CONVERGED) {
            delete out[i];
            out[i] = nullptr;
        }
    }
    if (n > 1 && share_lattices(out) != CTDSE_OK) {
        for (int i = 0; i < n; i++) {
            delete out[i];
            out[i] = nullptr;
            status[i] = g_status;
        }
        return g_status;
    }
    if (first != CTDSE_OK) set_status(first, first_msg);
    else clear_status();
    return first;
}

int geometry_from(int num_r, double dx, double dt, double a, double e_guess, double cv, double x_int,
                  const double *psi0_host, const double *x_host, ctdse_geometry **out)
{
    if (!out) return set_status(CTDSE_ERR_ARGUMENT, "null output pointer");
    *out = nullptr;
    std::vector<GeomSpec> spec{{num_r, dx, dt, a, e_guess, cv, x_int, psi0_host, x_host}};
    std::vector<ctdse_geometry *> gs;
    std::vector<int> st;
    const int code = geometries_from(spec, gs, st);
    *out = gs[0];
    return code;
}

// call1DTDSE's geometry cache: the last geometry it built, reused while the calls' grid,
// time step, atom, PopInt window, ground state and x grid are the same (a task farm of calls
// shares one ground state, MPI-TDSE.c:145), so a call after the first builds no tables.
// Held through shared_ptr: a call keeps its geometry alive while another call replaces the
// entry.  Never destroyed (device memory is not released from a static destructor).
struct GeomCache {
    std::mutex m;
    std::shared_ptr<ctdse_geometry> geom;
    int device = -1, num_r = 0;
    double dx = 0, dt = 0, a = 0, x_int = 0;
    std::vector<double> psi0, x;
};
GeomCache *g_geom_cache = new GeomCache;

std::shared_ptr<ctdse_geometry> cached_geometry(const inputs_def &in, double dt, int *st)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = -1;
    const size_t nx = in.num_r + 1;
    GeomCache &C = *g_geom_cache;
    std::lock_guard<std::mutex> lock(C.m);
    if (C.geom && C.device == dev && C.num_r == in.num_r && C.dx == in.dx && C.dt == dt && C.a == in.trg.a &&
        C.x_int == in.x_int && std::memcmp(C.psi0.data(), in.psi0, 2 * nx * sizeof(double)) == 0 &&
        std::memcmp(C.x.data(), in.x, nx * sizeof(double)) == 0) {
        *st = CTDSE_OK;
        return C.geom;
    }
    ctdse_geometry *G = nullptr;
    *st = geometry_from(in.num_r, in.dx, dt, in.trg.a, in.Eguess, in.CV, in.x_int, in.psi0, in.x, &G);
    if (*st != CTDSE_OK) {
        delete G;
        return nullptr;
    }
    C.geom.reset(G);
    C.device = dev;
    C.num_r = in.num_r;
    C.dx = in.dx;
    C.dt = dt;
    C.a = in.trg.a;
    C.x_int = in.x_int;
    C.psi0.assign(in.psi0, in.psi0 + 2 * nx);
    C.x.assign(in.x, in.x + nx);
    return C.geom;
}

struct InstSpec {
    const ctdse_geometry *geom;
    int n_steps;
    const double *field;     // device
    double *out;             // device, 5 rows of n_steps (may be null for partial runs)
    long long out_row = 0;   // distance between the 5 rows (0: n_steps, packed)
    double *snap = nullptr;  // device
    int snap_every = 1, n_snap = 0;
    double *psi_final = nullptr;
    int k0 = 0, n_run = 0, pregated = 0;
    const double2 *psi_init = nullptr;
    double2 *save = nullptr;
    int save_every = 1;
};

// Waves per workgroup.  A segment is nw*64 chunks of which 2H are truncation halo, so larger
// workgroups waste fewer lanes; but every segment starts with loads and every scan ends at a
// barrier, and those stalls are hidden only by other workgroups on the same CU in a different
// phase.  Measured on the 4096 x 16001-point batch (lane efficiency in brackets):
//   8 wave slots per CU (2 waves/SIMD): 4 x 2 WG/CU 3.00 s (0.78), 2 x 4 3.24 s, 3 x 2 3.54 s,
//                                       8 x 1 3.66 s, 7 x 1 3.68 s (0.89), 6 x 1 4.12 s;
//   12 wave slots (3 waves/SIMD):       4 x 3 2.51 s (0.78), 3 x 4 3.14 s, 2 x 6 4.32 s,
//                                       6 x 2 4.39 s (0.87), 5 x 2 5.39 s.
// Four-wave workgroups win in both settings, so they are used whenever the halo leaves at
// least one chunk per lane-row of a segment (H <= 96 lanes; the table build caps H lower).
int pick_waves(int *min_out, int max_halo_lanes)
{
    int min_nw = 1;
    while (min_nw * 64 - max_halo_lanes < 64) min_nw++;
    int nw = std::max(4, min_nw);
    if (const char *env = std::getenv("MMA_TDSE_WAVES")) nw = std::max(min_nw, std::min(TDSE_MAX_WAVES, std::atoi(env)));
    *min_out = min_nw;
    return nw;
}

// Launch shape of the propagation kernel: waves per workgroup (the segment is nw * 64
// chunks) and resident workgroups per CU (the persistent grid is CUs * per_cu).
struct LaunchShape {
    int nw = 4, per_cu = 3, steps = 1;
    bool stash = false;  // two-step passes keep 1/beta in LDS between their stages
};

// Dynamic LDS of tdse_propagate_pair_stash_kernel: two segments' 1/beta.
size_t stash_bytes(int nw) { return 2 * 64 * (size_t)nw * TDSE_CHUNK * sizeof(double2); }

// Whether per_cu workgroups of the stash kernel (static LDS + stash) fit one CU's LDS, so
// that the shape keeps its occupancy (MMA_TDSE_STASH=0 turns the stash off).
bool stash_fits(const LaunchShape &sh)
{
    if (const char *env = std::getenv("MMA_TDSE_STASH"))
        if (std::atoi(env) == 0) return false;
    hipFuncAttributes attr;
    int dev = 0, lds = 0;
    if (hipFuncGetAttributes(&attr, (const void *)tdse_propagate_pair_stash_kernel) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess)
        return false;
    return (size_t)sh.per_cu * (attr.sharedSizeBytes + stash_bytes(sh.nw)) <= (size_t)lds;
}

int shape_for(const std::vector<const ctdse_geometry *> &geoms, const ctdse_launch_config *cfg, LaunchShape *out)
{
    int min_nw = 1, max_h = 0;
    for (const ctdse_geometry *g : geoms) max_h = std::max(max_h, g->dev.halo_lanes);
    LaunchShape sh;
    sh.nw = pick_waves(&min_nw, max_h);
    sh.steps = kDefaultStepsPerPass;
    if (const char *env = std::getenv("MMA_TDSE_STEPS_PER_PASS")) sh.steps = std::atoi(env) == 2 ? 2 : 1;
    sh.per_cu = std::max(1, (sh.steps == 2 ? 4 * TDSE_PAIR_WAVES_PER_SIMD : TDSE_WAVES_PER_CU) / sh.nw);
    if (const char *env = std::getenv("MMA_TDSE_WG_PER_CU")) sh.per_cu = std::max(1, std::atoi(env));
    if (cfg) {
        if (cfg->waves_per_group < min_nw || cfg->waves_per_group > TDSE_MAX_WAVES)
            return set_status(CTDSE_ERR_ARGUMENT, "waves_per_group must lie in [" + std::to_string(min_nw) + ", " +
                                                      std::to_string(TDSE_MAX_WAVES) + "] for these geometries");
        if (cfg->groups_per_cu < 1 || cfg->groups_per_cu > 16)
            return set_status(CTDSE_ERR_ARGUMENT, "groups_per_cu must lie in [1, 16]");
        if (cfg->steps_per_pass != 1 && cfg->steps_per_pass != 2)
            return set_status(CTDSE_ERR_ARGUMENT, "steps_per_pass must be 1 or 2");
        sh.nw = cfg->waves_per_group;
        sh.per_cu = cfg->groups_per_cu;
        sh.steps = cfg->steps_per_pass;
    }
    sh.stash = sh.steps == 2 && stash_fits(sh);
    if (sh.stash)  // outside any stream capture: plans and immediate runs call this first
        HIP_CHECK(hipFuncSetAttribute((const void *)tdse_propagate_pair_stash_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)stash_bytes(sh.nw)));
    *out = sh;
    return CTDSE_OK;
}

} // namespace

// A prepared batch: device descriptors, the instance queue and a private work buffer, so
// that a launch is two kernel enqueues with no allocation, copy or synchronisation (and can
// be captured into a graph).
struct ctdse_batch_plan {
    int n = 0, slots = 0;
    LaunchShape shape;
    long long stride = 0;
    DevMem work, d_geom, d_inst, d_order, d_counter;
    TdseQueues queues{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;  // recorded after every run outside a capture
    ~ctdse_batch_plan()
    {
        // the last run may still be in flight: let it finish before its buffers go
        if (done) {
            (void)hipEventSynchronize(done);
            (void)hipEventDestroy(done);
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace {

// Descriptors of a batch on the host: validated instances, distinct geometries and the
// per-XCD queues of instances, each in longest-first order.
struct BatchHost {
    std::vector<const ctdse_geometry *> geoms;
    std::vector<TdseGeom> gdev;
    std::vector<TdseInst> insts;
    std::vector<int> order;
    TdseQueues queues{};
    long long stride = 0;
};

int describe(const std::vector<InstSpec> &spec, BatchHost &B)
{
    const int n = (int)spec.size();
    B.insts.assign(n, TdseInst{});
    for (int i = 0; i < n; i++) {
        const InstSpec &s = spec[i];
        if (!s.geom || s.n_steps < 1 || !s.field || (!s.out && !s.save && s.n_run == 0) || s.k0 < 0 || s.n_run < 0 ||
            s.k0 + s.n_run > s.n_steps || (s.k0 > 0 && !s.pregated) || (s.save && s.save_every < 1))
            return set_status(CTDSE_ERR_ARGUMENT, "instance " + std::to_string(i) + ": bad descriptor");
        int gi = (int)(std::find(B.geoms.begin(), B.geoms.end(), s.geom) - B.geoms.begin());
        if (gi == (int)B.geoms.size()) B.geoms.push_back(s.geom);
        TdseInst &I = B.insts[i];
        I.geom = gi;
        I.n_steps = s.n_steps;
        I.field = s.field;
        const long long row = s.out_row > 0 ? s.out_row : s.n_steps;
        if (s.out_row > 0 && s.out_row < s.n_steps)
            return set_status(CTDSE_ERR_ARGUMENT, "instance " + std::to_string(i) + ": output rows shorter than n_steps");
        I.src = s.out;
        I.ptot = s.out ? s.out + row : nullptr;
        I.pint = s.out ? s.out + 2 * row : nullptr;
        I.expv = s.out ? s.out + 3 * row : nullptr;
        I.efield = s.out ? s.out + 4 * row : nullptr;
        I.k0 = s.k0;
        I.n_run = s.n_run;
        I.pregated = s.pregated;
        I.psi_init = s.psi_init;
        I.save = s.save;
        I.save_every = s.save_every;
        // a state vector given as both start and save slot of a run of save_every steps is
        // stepped in place
        I.inplace = s.psi_init && (const void *)s.psi_init == (const void *)s.save && s.n_run == s.save_every;
        I.snap = s.snap;
        I.snap_every = std::max(1, s.snap_every);
        I.n_snap = s.n_snap;
        I.psi_final = s.psi_final;
        B.stride = std::max(B.stride, s.geom->dev.npad);
    }
    // largest instances first (longest-processing-time order for the atomic queues)
    std::vector<int> lpt(n);
    for (int i = 0; i < n; i++) lpt[i] = i;
    auto steps_of = [&](int p) { return (double)(spec[p].n_run > 0 ? spec[p].n_run : spec[p].n_steps - spec[p].k0); };
    std::stable_sort(lpt.begin(), lpt.end(), [&](int p, int q) {
        return (double)spec[p].geom->dev.nx * steps_of(p) > (double)spec[q].geom->dev.nx * steps_of(q);
    });
    // one queue per XCD: an instance whose geometry reads shared lattice tables goes to the
    // queue of its alignment (the 8 alignments of a lattice are 8 sets of lines; an XCD's 4 MB
    // L2 then holds one of them), the others are dealt round-robin
    static_assert(Lattice::kAlign == TDSE_QUEUES, "one queue per lattice alignment");
    std::vector<int> per_queue[TDSE_QUEUES];
    int dealt = 0;
    for (int p : lpt) {
        const int shift = spec[p].geom->lattice ? spec[p].geom->lattice_shift : -1;
        per_queue[shift >= 0 ? shift : dealt++ % TDSE_QUEUES].push_back(p);
    }
    B.order.clear();
    B.queues.off[0] = 0;
    for (int q = 0; q < TDSE_QUEUES; q++) {
        B.order.insert(B.order.end(), per_queue[q].begin(), per_queue[q].end());
        B.queues.off[q + 1] = (int)B.order.size();
    }
    B.gdev.resize(B.geoms.size());
    for (size_t k = 0; k < B.geoms.size(); k++) B.gdev[k] = B.geoms[k]->dev;
    return CTDSE_OK;
}

// Persistent workgroups of a launch: CUs * per_cu, or fewer for a smaller batch.  (Spreading
// a batch over equal rounds instead, e.g. 683 workgroups for 4096 instances, measured slower:
// 2.04 s against 1.95 s; the last, partly filled round runs faster per instance.)
int grid_slots(int n, const LaunchShape &sh, int *slots)
{
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    int cus = 0;
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    *slots = std::max(1, std::min(n, cus * sh.per_cu));
    return CTDSE_OK;
}

// Enqueues the queue reset and the persistent kernel (no allocation, copy or wait).
int enqueue_plan(const ctdse_batch_plan &P, double2 *work, hipStream_t stream)
{
    hipLaunchKernelGGL(tdse_fill_int_kernel, dim3(1), dim3(64), 0, stream, P.d_counter.as<int>(), TDSE_QUEUES, 0);
    const bool stash = P.shape.steps == 2 && P.shape.stash;
    hipLaunchKernelGGL(stash ? tdse_propagate_pair_stash_kernel
                             : (P.shape.steps == 2 ? tdse_propagate_pair_kernel : tdse_propagate_kernel),
                       dim3(P.slots), dim3(64 * P.shape.nw), stash ? stash_bytes(P.shape.nw) : 0, stream,
                       P.d_geom.as<TdseGeom>(), P.d_inst.as<TdseInst>(), P.d_order.as<int>(), P.queues,
                       P.d_counter.as<int>(), work, P.stride);
    HIP_CHECK(hipGetLastError());
    return CTDSE_OK;
}

// Immediate path: descriptors copied on `stream`, the shared work buffer, and a wait.
int propagate(const std::vector<InstSpec> &spec, hipStream_t stream, const ctdse_launch_config *cfg = nullptr)
{
    const int n = (int)spec.size();
    if (n == 0) return CTDSE_OK;
    BatchHost B;
    if (describe(spec, B) != CTDSE_OK) return g_status;
    ctdse_batch_plan P;
    P.n = n;
    P.queues = B.queues;
    P.stride = B.stride;
    if (shape_for(B.geoms, cfg, &P.shape) != CTDSE_OK) return g_status;
    if (grid_slots(n, P.shape, &P.slots) != CTDSE_OK) return g_status;

    std::lock_guard<std::mutex> lock(g_launch_mutex);
    const size_t need = (size_t)P.slots * P.stride * sizeof(double2);
    if (need > g_work_bytes) {
        HIP_CHECK(g_work.alloc(need));
        g_work_bytes = need;
    }
    HIP_CHECK(P.d_geom.alloc(B.gdev.size() * sizeof(TdseGeom)));
    HIP_CHECK(P.d_inst.alloc(B.insts.size() * sizeof(TdseInst)));
    HIP_CHECK(P.d_order.alloc(B.order.size() * sizeof(int)));
    HIP_CHECK(P.d_counter.alloc(TDSE_QUEUES * sizeof(int)));
    HIP_CHECK(hipMemcpyAsync(P.d_geom.p, B.gdev.data(), B.gdev.size() * sizeof(TdseGeom), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(P.d_inst.p, B.insts.data(), B.insts.size() * sizeof(TdseInst), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(P.d_order.p, B.order.data(), B.order.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    const int queued = enqueue_plan(P, g_work.as<double2>(), stream);
    HIP_CHECK(hipStreamSynchronize(stream));  // the descriptors are freed on return
    return queued;
}

// Waves per workgroup of the step-parallel path (MMA_TDSE_PAR_WAVES overrides): segments of
// nw*64 - 2H own chunks, one workgroup each, one launch per step.
int par_waves(int halo_lanes)
{
    int nw = 4;
    if (const char *env = std::getenv("MMA_TDSE_PAR_WAVES")) nw = std::atoi(env);
    nw = std::max(1, std::min(TDSE_MAX_WAVES, nw));
    while (nw < TDSE_MAX_WAVES && nw * 64 - 2 * halo_lanes < 64) nw++;
    return nw * 64 - 2 * halo_lanes >= 64 ? nw : 0;
}

// Descriptors of a step-parallel run written on the device from kernel arguments, so the
// run needs no host-to-device copy (and no staging memory).
__global__ void tdse_put_desc_kernel(TdseGeom g, TdseInst inst, TdseGeom *g_out, TdseInst *inst_out)
{
    if (threadIdx.x == 0) {
        *g_out = g;
        *inst_out = inst;
    }
}

// Launch shape of a step-parallel run of one instance (propagate_steps): waves per
// workgroup and segments per step.  Returns CTDSE_ERR_UNSUPPORTED (status untouched) when the
// descriptor needs the persistent path.
int steps_shape(const InstSpec &s, int *nw_out, int *nseg_out)
{
    if (!s.geom || s.n_run != 0 || s.k0 != 0 || s.psi_init || s.save || s.pregated || !s.out || s.n_steps < 1)
        return CTDSE_ERR_UNSUPPORTED;
    const TdseGeom &g = s.geom->dev;
    const int nw = par_waves(g.halo_lanes);
    if (nw == 0) return CTDSE_ERR_UNSUPPORTED;
    const int own = nw * 64 - 2 * g.halo_lanes;
    *nw_out = nw;
    *nseg_out = (g.nchunks + own - 1) / own;
    return CTDSE_OK;
}

// Device buffers of a step-parallel run: the two descriptors, per-segment partial sums of
// the observables (2 x nseg x 5 doubles), the gated field (n_steps + 1 doubles) and two
// state vectors (npad complex doubles each).
struct StepBufs {
    TdseGeom *geom;
    TdseInst *inst;
    double *partial;
    double *gated;
    double2 *psi[2];
};

// Queues a step-parallel run into the given buffers: no allocation and no wait.  Every
// buffer is (re)initialised here, so buffers left over from an earlier run may be reused.
int enqueue_steps(const InstSpec &s, int nw, int nseg, const StepBufs &b, hipStream_t stream)
{
    const TdseGeom &g = s.geom->dev;
    BatchHost B;
    if (describe({s}, B) != CTDSE_OK) return g_status;
    const int n_field = s.n_steps + 1;
    TdseInst inst = B.insts[0];
    inst.field = b.gated;
    inst.pregated = 1;
    hipLaunchKernelGGL(tdse_put_desc_kernel, dim3(1), dim3(64), 0, stream, g, inst, b.geom, b.inst);
    // psi0 into the first buffer; the second one zeroed (by a kernel, not a memset node), its
    // zero tiles and padding rows stay zero (the steps write own chunks only)
    HIP_CHECK(hipMemcpyAsync(b.psi[0], g.psi0, (size_t)g.npad * sizeof(double2), hipMemcpyDeviceToDevice, stream));
    const long long n_int = 4 * g.npad;
    hipLaunchKernelGGL(tdse_fill_int_kernel, dim3((unsigned)((n_int + 255) / 256)), dim3(256), 0, stream,
                       reinterpret_cast<int *>(b.psi[1]), (int)n_int, 0);
    hipLaunchKernelGGL(tdse_gate_kernel, dim3(1), dim3(256), 0, stream, s.field, n_field, b.gated);
    const dim3 untile_grid((g.nx + 255) / 256);
    const int n_snap = s.snap ? s.n_snap : 0, every = std::max(1, s.snap_every);
    int i_snap = 0;
    for (int k = 0; k < s.n_steps; k++) {
        hipLaunchKernelGGL(tdse_parallel_step_kernel, dim3(nseg, 1), dim3(64 * nw), 0, stream, b.geom, b.inst, k,
                           b.psi[k & 1], b.psi[(k + 1) & 1], (long long)g.npad, b.partial, nseg);
        // prop.c:248-255: row i_snap gets psi after step k when k % every == 0
        if (k % every == 0 && i_snap < n_snap - 1)
            hipLaunchKernelGGL(tdse_untile_kernel, untile_grid, dim3(256), 0, stream, b.psi[(k + 1) & 1], g.nx,
                               s.snap + (size_t)i_snap++ * 2 * g.nx);
    }
    // prop.c:256-262: after the last step, the last row gets the final state if every earlier
    // row was used; the final state itself is optional output
    const double2 *fin = b.psi[s.n_steps & 1];
    if (n_snap > 0 && i_snap == n_snap - 1)
        hipLaunchKernelGGL(tdse_untile_kernel, untile_grid, dim3(256), 0, stream, fin, g.nx,
                           s.snap + (size_t)(n_snap - 1) * 2 * g.nx);
    if (s.psi_final) hipLaunchKernelGGL(tdse_untile_kernel, untile_grid, dim3(256), 0, stream, fin, g.nx, s.psi_final);
    hipLaunchKernelGGL(tdse_parallel_finish_kernel, dim3(1), dim3(64), 0, stream, b.inst, 1, b.partial, nseg);
    HIP_CHECK(hipGetLastError());
    return CTDSE_OK;
}

// One instance over its whole run with no partial runs or saved states: the segments of a
// step are solved side by side (tdse_parallel_step_kernel, psi double-buffered), so a step
// costs one segment's latency instead of nseg of them; snapshots and the final state are
// copied out between steps.  Returns CTDSE_ERR_UNSUPPORTED (status untouched) when the
// descriptor needs the persistent path.
int propagate_steps(const InstSpec &s, hipStream_t stream)
{
    int nw = 0, nseg = 0;
    const int shaped = steps_shape(s, &nw, &nseg);
    if (shaped != CTDSE_OK) return shaped;
    const TdseGeom &g = s.geom->dev;
    DevMem states, partial, gated, d_geom, d_inst;
    HIP_CHECK(states.alloc(2 * (size_t)g.npad * sizeof(double2)));
    HIP_CHECK(partial.alloc(2 * (size_t)nseg * 5 * sizeof(double)));
    HIP_CHECK(gated.alloc((size_t)(s.n_steps + 1) * sizeof(double)));
    HIP_CHECK(d_geom.alloc(sizeof(TdseGeom)));
    HIP_CHECK(d_inst.alloc(sizeof(TdseInst)));
    const StepBufs b{d_geom.as<TdseGeom>(), d_inst.as<TdseInst>(), partial.as<double>(), gated.as<double>(),
                     {states.as<double2>(), states.as<double2>() + g.npad}};
    std::lock_guard<std::mutex> lock(g_launch_mutex);
    const int queued = enqueue_steps(s, nw, nseg, b, stream);
    // wait before returning (also after a launch error): the buffers freed on return may
    // still be in use
    HIP_CHECK(hipStreamSynchronize(stream));
    return queued;
}

// The device address of this GPU's `scratch`, looked up once per device with that device
// current (each GPU holds its own copy of the array at its own address).
std::mutex g_scratch_mutex;
constexpr int kMaxDevices = 64;
void *g_scratch_addr[kMaxDevices] = {};

int scratch_address(void **out)
{
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_status(CTDSE_ERR_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_status(CTDSE_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return CTDSE_OK;
}

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// Plan path: everything allocated and uploaded here, once.
int make_plan(const std::vector<InstSpec> &spec, const ctdse_launch_config *cfg, ctdse_batch_plan &P)
{
    const int n = (int)spec.size();
    BatchHost B;
    if (describe(spec, B) != CTDSE_OK) return g_status;
    P.n = n;
    P.queues = B.queues;
    P.stride = B.stride;
    if (shape_for(B.geoms, cfg, &P.shape) != CTDSE_OK) return g_status;
    if (grid_slots(std::max(n, 1), P.shape, &P.slots) != CTDSE_OK) return g_status;
    HIP_CHECK(P.work.alloc((size_t)P.slots * P.stride * sizeof(double2)));
    HIP_CHECK(P.d_geom.alloc(B.gdev.size() * sizeof(TdseGeom)));
    HIP_CHECK(P.d_inst.alloc(B.insts.size() * sizeof(TdseInst)));
    HIP_CHECK(P.d_order.alloc(B.order.size() * sizeof(int)));
    HIP_CHECK(P.d_counter.alloc(TDSE_QUEUES * sizeof(int)));
    HIP_CHECK(hipMemcpy(P.d_geom.p, B.gdev.data(), B.gdev.size() * sizeof(TdseGeom), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(P.d_inst.p, B.insts.data(), B.insts.size() * sizeof(TdseInst), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(P.d_order.p, B.order.data(), B.order.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_CHECK(hipEventCreateWithFlags(&P.done, hipEventDisableTiming));
    return CTDSE_OK;
}

// Marks the end of a run on `stream` for the plan's destructor (not while the stream is being
// captured: there the caller's graph owns the work).
int mark_done(ctdse_batch_plan &P, hipStream_t stream)
{
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIP_CHECK(hipStreamIsCapturing(stream, &cap));
    if (cap == hipStreamCaptureStatusNone && P.done) HIP_CHECK(hipEventRecord(P.done, stream));
    return CTDSE_OK;
}

int adjoint(const std::vector<const ctdse_geometry *> &ig, const std::vector<TdseAdjInst> &ai, hipStream_t stream)
{
    const int n = (int)ai.size();
    if (n == 0) return CTDSE_OK;
    std::vector<const ctdse_geometry *> geoms;
    std::vector<TdseAdjInst> insts = ai;
    std::vector<double> work(n);
    for (int i = 0; i < n; i++) {
        int gi = (int)(std::find(geoms.begin(), geoms.end(), ig[i]) - geoms.begin());
        if (gi == (int)geoms.size()) geoms.push_back(ig[i]);
        insts[i].geom = gi;
        work[i] = (double)(ai[i].k_hi - ai[i].k_lo);
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) {
        return (double)ig[p]->dev.nx * work[p] > (double)ig[q]->dev.nx * work[q];
    });
    std::vector<TdseGeom> gdev(geoms.size());
    for (size_t k = 0; k < geoms.size(); k++) gdev[k] = geoms[k]->dev;
    int min_nw = 1, max_h = 0;
    for (const ctdse_geometry *g : geoms) max_h = std::max(max_h, g->dev.halo_lanes);
    const int nw = pick_waves(&min_nw, max_h);
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    const int per_cu = std::max(1, TDSE_ADJ_WAVES_PER_CU / nw);
    const int slots = std::max(1, std::min(n, prop.multiProcessorCount * per_cu));
    std::lock_guard<std::mutex> lock(g_launch_mutex);
    DevMem d_geom, d_inst, d_order, d_counter;
    HIP_CHECK(d_geom.alloc(gdev.size() * sizeof(TdseGeom)));
    HIP_CHECK(d_inst.alloc(insts.size() * sizeof(TdseAdjInst)));
    HIP_CHECK(d_order.alloc(order.size() * sizeof(int)));
    HIP_CHECK(d_counter.alloc(sizeof(int)));
    HIP_CHECK(hipMemcpyAsync(d_geom.p, gdev.data(), gdev.size() * sizeof(TdseGeom), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(d_inst.p, insts.data(), insts.size() * sizeof(TdseAdjInst), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(d_order.p, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(tdse_fill_int_kernel, dim3(1), dim3(64), 0, stream, d_counter.as<int>(), 1, 0);
    hipLaunchKernelGGL(tdse_adjoint_kernel, dim3(slots), dim3(64 * nw), 0, stream, d_geom.as<TdseGeom>(),
                       d_inst.as<TdseAdjInst>(), d_order.as<int>(), n, d_counter.as<int>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(stream));
    return CTDSE_OK;
}

int reverse_adjoint(const std::vector<const ctdse_geometry *> &ig, const std::vector<TdseRevInst> &ri,
                    hipStream_t stream)
{
    const int n = (int)ri.size();
    if (n == 0) return CTDSE_OK;
    std::vector<const ctdse_geometry *> geoms;
    std::vector<TdseRevInst> insts = ri;
    long long stride = 0;
    for (int i = 0; i < n; i++) {
        int gi = (int)(std::find(geoms.begin(), geoms.end(), ig[i]) - geoms.begin());
        if (gi == (int)geoms.size()) geoms.push_back(ig[i]);
        insts[i].geom = gi;
        stride = std::max(stride, ig[i]->dev.npad);
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) {
        return (double)ig[p]->dev.nx * ri[p].k_hi > (double)ig[q]->dev.nx * ri[q].k_hi;
    });
    std::vector<TdseGeom> gdev(geoms.size());
    for (size_t k = 0; k < geoms.size(); k++) gdev[k] = geoms[k]->dev;
    int min_nw = 1, max_h = 0;
    for (const ctdse_geometry *g : geoms) max_h = std::max(max_h, g->dev.halo_lanes);
    const int nw = pick_waves(&min_nw, max_h);
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    const int per_cu = std::max(1, 4 * TDSE_REV_WAVES_PER_SIMD / nw);
    const int slots = std::max(1, std::min(n, prop.multiProcessorCount * per_cu));
    std::lock_guard<std::mutex> lock(g_launch_mutex);
    const size_t need = (size_t)slots * 2 * stride * sizeof(double2);  // psi and kappa per workgroup
    if (need > g_work_bytes) {
        HIP_CHECK(g_work.alloc(need));
        g_work_bytes = need;
    }
    DevMem d_geom, d_inst, d_order, d_counter;
    HIP_CHECK(d_geom.alloc(gdev.size() * sizeof(TdseGeom)));
    HIP_CHECK(d_inst.alloc(insts.size() * sizeof(TdseRevInst)));
    HIP_CHECK(d_order.alloc(order.size() * sizeof(int)));
    HIP_CHECK(d_counter.alloc(sizeof(int)));
    HIP_CHECK(hipMemcpyAsync(d_geom.p, gdev.data(), gdev.size() * sizeof(TdseGeom), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(d_inst.p, insts.data(), insts.size() * sizeof(TdseRevInst), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(d_order.p, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(tdse_fill_int_kernel, dim3(1), dim3(64), 0, stream, d_counter.as<int>(), 1, 0);
    const size_t park = (size_t)nw * TDSE_TILE * sizeof(double2);  // one parked state chunk per lane
    hipLaunchKernelGGL(tdse_reverse_adjoint_kernel, dim3(slots), dim3(64 * nw), park, stream, d_geom.as<TdseGeom>(),
                       d_inst.as<TdseRevInst>(), d_order.as<int>(), n, d_counter.as<int>(), g_work.as<double2>(),
                       stride);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(stream));
    return CTDSE_OK;
}

double *host_copy(const double *dev, size_t n, size_t alloc_n)
{
    double *h = (double *)calloc(std::max<size_t>(alloc_n, 1), sizeof(double));
    if (h && n && hipMemcpy(h, dev, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        free(h);
        return nullptr;
    }
    return h;
}

int spectrum_to_host(const double *sig_dev, int n, double dt, double tmax, double **omega, double **F, double **M2)
{
    const int nc = n / 2 + 1;
    DevMem X, om, f, m2;
    HIP_CHECK(X.alloc(nc * sizeof(double2)));
    HIP_CHECK(om.alloc(nc * sizeof(double)));
    HIP_CHECK(f.alloc(2 * nc * sizeof(double)));
    HIP_CHECK(m2.alloc(nc * sizeof(double)));
    hipLaunchKernelGGL(tdse_dft_r2c_kernel, dim3((nc + 127) / 128), dim3(128), 0, 0, sig_dev, n, nc, X.as<double2>());
    hipLaunchKernelGGL(tdse_spectrum_kernel, dim3((nc + 255) / 256), dim3(256), 0, 0, X.as<double2>(), nc, dt, tmax,
                       om.as<double>(), f.as<double>(), m2.as<double>());
    HIP_CHECK(hipGetLastError());
    if (omega) *omega = host_copy(om.as<double>(), nc, nc);
    *F = host_copy(f.as<double>(), 2 * nc, 2 * nc);
    *M2 = host_copy(m2.as<double>(), nc, nc);
    if ((omega && !*omega) || !*F || !*M2)
        return set_status(CTDSE_ERR_DEVICE, "spectrum: allocating or copying the host arrays failed");
    return CTDSE_OK;
}

} // namespace

extern "C" {

// status hook for the analysis entry points (tdse_analysis.hip); not exported
__attribute__((visibility("hidden"))) int ctdse_set_status_from_analysis(int code, const char *msg)
{
    return set_status(code, msg ? msg : "");
}

int ctdse_last_error(void) { return g_status; }
const char *ctdse_last_error_message(void) { return g_message.c_str(); }

int ctdse_geometry_create(int num_r, double dx, double dt, double a, double e_guess, double cv, double x_int,
                          const double *psi0, ctdse_geometry **out)
{
    clear_status();
    return geometry_from(num_r, dx, dt, a, e_guess, cv, x_int, psi0, nullptr, out);
}

int ctdse_geometry_info(const ctdse_geometry *g, double *e_init, int *gs_iterations, int *halo_points)
{
    clear_status();
    if (!g) return set_status(CTDSE_ERR_ARGUMENT, "null geometry");
    if (e_init) *e_init = g->e_init;
    if (gs_iterations) *gs_iterations = g->gs_iters;
    if (halo_points) *halo_points = g->dev.halo_lanes * TDSE_CHUNK;
    return CTDSE_OK;
}

int ctdse_geometry_ground_state(const ctdse_geometry *g, double *x, double *psi0)
{
    clear_status();
    if (!g) return set_status(CTDSE_ERR_ARGUMENT, "null geometry");
    const size_t nx = g->num_r + 1;
    if (x) HIP_CHECK(hipMemcpy(x, g->x.p, nx * sizeof(double), hipMemcpyDeviceToHost));
    if (psi0) HIP_CHECK(hipMemcpy(psi0, g->psi0_nat.p, 2 * nx * sizeof(double), hipMemcpyDeviceToHost));
    return CTDSE_OK;
}

void ctdse_geometry_destroy(ctdse_geometry *g) { delete g; }

int ctdse_propagate_batch(int n_inst, const ctdse_geometry *const *geom, const int *n_steps, const double *field,
                          const long long *field_off, double *out, const long long *out_off, void *stream)
{
    clear_status();
    if (n_inst < 0 || (n_inst > 0 && (!geom || !n_steps || !field || !field_off || !out || !out_off)))
        return set_status(CTDSE_ERR_ARGUMENT, "null batch argument");
    if (device_ready() != CTDSE_OK) return g_status;
    std::vector<InstSpec> spec(n_inst);
    for (int i = 0; i < n_inst; i++) {
        spec[i].geom = geom[i];
        spec[i].n_steps = n_steps[i];
        spec[i].field = field + field_off[i];
        spec[i].out = out + out_off[i];
    }
    return propagate(spec, (hipStream_t)stream);
}

int ctdse_propagate_one(const ctdse_geometry *geom, int n_steps, const double *field, double *out, void *stream)
{
    clear_status();
    if (!geom || !field || !out || n_steps < 1) return set_status(CTDSE_ERR_ARGUMENT, "null or empty argument");
    if (device_ready() != CTDSE_OK) return g_status;
    InstSpec s;
    s.geom = geom;
    s.n_steps = n_steps;
    s.field = field;
    s.out = out;
    const int st = propagate_steps(s, (hipStream_t)stream);
    if (st != CTDSE_ERR_UNSUPPORTED) return st;
    const ctdse_launch_config single{TDSE_MAX_WAVES, 1, 1};
    return propagate({s}, (hipStream_t)stream, &single);
}

int ctdse_propagate_one_noalloc(const ctdse_geometry *geom, int n_steps, const double *field, double *out,
                                void *stream)
{
    clear_status();
    if (!geom || !field || !out || n_steps < 1) return set_status(CTDSE_ERR_ARGUMENT, "null or empty argument");
    if (device_ready() != CTDSE_OK) return g_status;
    InstSpec s;
    s.geom = geom;
    s.n_steps = n_steps;
    s.field = field;
    s.out = out;
    int nw = 0, nseg = 0;
    if (steps_shape(s, &nw, &nseg) != CTDSE_OK)
        return set_status(CTDSE_ERR_UNSUPPORTED, "the geometry's truncation halo leaves no room for the step-parallel "
                                                 "solve, which is the only form without a work buffer");
    const TdseGeom &g = geom->dev;
    // layout of the scratch array: descriptors, partial sums, gated field, two states
    size_t off = 0;
    const size_t o_geom = off;
    off += align256(sizeof(TdseGeom));
    const size_t o_inst = off;
    off += align256(sizeof(TdseInst));
    const size_t o_part = off;
    off += align256(2 * (size_t)nseg * 5 * sizeof(double));
    const size_t o_gated = off;
    off += align256((size_t)(n_steps + 1) * sizeof(double));
    const size_t o_psi = off;
    off += 2 * (size_t)g.npad * sizeof(double2);
    if (off > (size_t)CTDSE_NOALLOC_SCRATCH_BYTES)
        return set_status(CTDSE_ERR_ARGUMENT, "ctdse_propagate_one_noalloc: this run needs " + std::to_string(off) +
                                                  " bytes of scratch (two states of " + std::to_string(g.npad) +
                                                  " complex doubles, " + std::to_string(n_steps + 1) +
                                                  " field samples), more than the " +
                                                  std::to_string((size_t)CTDSE_NOALLOC_SCRATCH_BYTES) + " it holds");
    void *base = nullptr;
    if (scratch_address(&base) != CTDSE_OK) return g_status;
    char *p = static_cast<char *>(base);
    double2 *psi = reinterpret_cast<double2 *>(p + o_psi);
    const StepBufs b{reinterpret_cast<TdseGeom *>(p + o_geom), reinterpret_cast<TdseInst *>(p + o_inst),
                     reinterpret_cast<double *>(p + o_part), reinterpret_cast<double *>(p + o_gated),
                     {psi, psi + g.npad}};
    std::lock_guard<std::mutex> lock(g_launch_mutex);
    return enqueue_steps(s, nw, nseg, b, (hipStream_t)stream);
}

int ctdse_geometry_create_batch(int n, const int *num_r, const double *dx, const double *dt, const double *a,
                                const double *e_guess, const double *cv, const double *x_int, ctdse_geometry **out,
                                int *status)
{
    clear_status();
    if (n < 0 || (n > 0 && (!num_r || !dx || !dt || !a || !e_guess || !cv || !x_int || !out)))
        return set_status(CTDSE_ERR_ARGUMENT, "null geometry batch argument");
    std::vector<GeomSpec> spec(n);
    for (int i = 0; i < n; i++) spec[i] = GeomSpec{num_r[i], dx[i], dt[i], a[i], e_guess[i], cv[i], x_int[i], nullptr, nullptr};
    std::vector<ctdse_geometry *> gs;
    std::vector<int> st;
    const int code = geometries_from(spec, gs, st);
    for (int i = 0; i < n; i++) {
        out[i] = gs.empty() ? nullptr : gs[i];
        if (status) status[i] = st.empty() ? code : st[i];
    }
    return code;
}

int ctdse_launch_config_default(int n_inst, const ctdse_geometry *const *geom, ctdse_launch_config *cfg)
{
    clear_status();
    if (n_inst < 1 || !geom || !cfg) return set_status(CTDSE_ERR_ARGUMENT, "null argument");
    std::vector<const ctdse_geometry *> gs;
    for (int i = 0; i < n_inst; i++) {
        if (!geom[i]) return set_status(CTDSE_ERR_ARGUMENT, "null geometry " + std::to_string(i));
        if (std::find(gs.begin(), gs.end(), geom[i]) == gs.end()) gs.push_back(geom[i]);
    }
    LaunchShape sh;
    if (shape_for(gs, nullptr, &sh) != CTDSE_OK) return g_status;
    cfg->waves_per_group = sh.nw;
    cfg->groups_per_cu = sh.per_cu;
    cfg->steps_per_pass = sh.steps;
    return CTDSE_OK;
}

int ctdse_batch_plan_create(int n_inst, const ctdse_geometry *const *geom, const int *n_steps, const double *field,
                            const long long *field_off, double *out, const long long *out_off,
                            long long out_row_stride, const ctdse_launch_config *config, ctdse_batch_plan **plan)
{
    clear_status();
    if (!plan) return set_status(CTDSE_ERR_ARGUMENT, "null plan pointer");
    *plan = nullptr;
    if (n_inst < 1 || !geom || !n_steps || !field || !field_off || !out || !out_off)
        return set_status(CTDSE_ERR_ARGUMENT, "null batch argument");
    if (out_row_stride < 0) return set_status(CTDSE_ERR_ARGUMENT, "negative output row stride");
    if (device_ready() != CTDSE_OK) return g_status;
    std::vector<InstSpec> spec(n_inst);
    for (int i = 0; i < n_inst; i++) {
        spec[i].geom = geom[i];
        spec[i].n_steps = n_steps[i];
        spec[i].field = field + field_off[i];
        spec[i].out = out + out_off[i];
        spec[i].out_row = out_row_stride;
    }
    auto *P = new ctdse_batch_plan;
    if (make_plan(spec, config, *P) != CTDSE_OK) {
        delete P;
        return g_status;
    }
    *plan = P;
    return CTDSE_OK;
}

int ctdse_batch_plan_enqueue(ctdse_batch_plan *plan, void *stream)
{
    clear_status();
    if (!plan) return set_status(CTDSE_ERR_ARGUMENT, "null plan");
    if (enqueue_plan(*plan, plan->work.as<double2>(), (hipStream_t)stream) != CTDSE_OK) return g_status;
    return mark_done(*plan, (hipStream_t)stream);
}

int ctdse_batch_plan_graph_launch(ctdse_batch_plan *plan, void *stream)
{
    clear_status();
    if (!plan) return set_status(CTDSE_ERR_ARGUMENT, "null plan");
    if (!plan->exec) {
        hipStream_t cs = nullptr;
        HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        int st = CTDSE_OK;
        if (e == hipSuccess) st = enqueue_plan(*plan, plan->work.as<double2>(), cs);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(cs, &g);
        (void)hipStreamDestroy(cs);
        if (e != hipSuccess || e2 != hipSuccess || st != CTDSE_OK) {
            if (g) (void)hipGraphDestroy(g);
            if (st != CTDSE_OK) return st;
            return set_status(CTDSE_ERR_DEVICE, std::string("graph capture: ") +
                                                    hipGetErrorString(e != hipSuccess ? e : e2));
        }
        plan->graph = g;
        HIP_CHECK(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
    }
    HIP_CHECK(hipGraphLaunch(plan->exec, (hipStream_t)stream));
    return mark_done(*plan, (hipStream_t)stream);
}

int ctdse_batch_plan_info(const ctdse_batch_plan *plan, ctdse_launch_config *config, int *workgroups)
{
    clear_status();
    if (!plan) return set_status(CTDSE_ERR_ARGUMENT, "null plan");
    if (config) {
        config->waves_per_group = plan->shape.nw;
        config->groups_per_cu = plan->shape.per_cu;
        config->steps_per_pass = plan->shape.steps;
    }
    if (workgroups) *workgroups = plan->slots;
    return CTDSE_OK;
}

void ctdse_batch_plan_destroy(ctdse_batch_plan *plan) { delete plan; }

long long ctdse_state_length(const ctdse_geometry *g)
{
    clear_status();
    if (!g) {
        set_status(CTDSE_ERR_ARGUMENT, "null geometry");
        return -1;
    }
    return g->dev.npad;
}

int ctdse_propagate_states(int n_inst, const ctdse_geometry *const *geom, const int *n_steps, const int *k0,
                           const int *n_run, const double *field, const long long *field_off, const double *psi_init,
                           const long long *init_off, double *save, const long long *save_off, int save_every,
                           double *out, const long long *out_off, void *stream)
{
    clear_status();
    if (n_inst < 0 || (n_inst > 0 && (!geom || !n_steps || !k0 || !n_run || !field || !field_off)))
        return set_status(CTDSE_ERR_ARGUMENT, "null batch argument");
    if (device_ready() != CTDSE_OK) return g_status;
    std::vector<InstSpec> spec(n_inst);
    for (int i = 0; i < n_inst; i++) {
        InstSpec &t = spec[i];
        t.geom = geom[i];
        t.n_steps = n_steps[i];
        t.k0 = k0[i];
        t.n_run = n_run[i];
        t.pregated = 1;
        t.field = field + field_off[i];
        t.psi_init = psi_init ? (const double2 *)(psi_init + init_off[i]) : nullptr;
        t.save = save ? (double2 *)(save + save_off[i]) : nullptr;
        t.save_every = save_every;
        t.out = out ? out + out_off[i] : nullptr;
    }
    return propagate(spec, (hipStream_t)stream);
}

int ctdse_adjoint_batch(int n_inst, const ctdse_geometry *const *geom, const int *k_hi, const int *k_lo,
                        const double *field, const long long *field_off, const double *gobs, const long long *gobs_off,
                        const double *states, const long long *states_off, double *kappa, const long long *kappa_off,
                        double *grad, const long long *grad_off, void *stream)
{
    clear_status();
    if (n_inst < 0 || (n_inst > 0 && (!geom || !k_hi || !k_lo || !field || !field_off || !gobs || !gobs_off ||
                                      !states || !states_off || !kappa || !kappa_off || !grad || !grad_off)))
        return set_status(CTDSE_ERR_ARGUMENT, "null adjoint argument");
    if (device_ready() != CTDSE_OK) return g_status;
    std::vector<const ctdse_geometry *> ig(n_inst);
    std::vector<TdseAdjInst> ai(n_inst);
    for (int i = 0; i < n_inst; i++) {
        if (!geom[i] || k_hi[i] <= k_lo[i] || k_lo[i] < 0)
            return set_status(CTDSE_ERR_ARGUMENT, "adjoint instance " + std::to_string(i) + ": bad range");
        ig[i] = geom[i];
        TdseAdjInst &a = ai[i];
        a.k_hi = k_hi[i];
        a.k_lo = k_lo[i];
        a.field = field + field_off[i];
        a.gobs = gobs + gobs_off[i];
        a.psi = (const double2 *)(states + states_off[i]);
        a.kappa = (double2 *)(kappa + kappa_off[i]);
        a.grad = grad + grad_off[i];
    }
    return adjoint(ig, ai, (hipStream_t)stream);
}

int ctdse_adjoint_reverse_batch(int n_inst, const ctdse_geometry *const *geom, const int *k_hi, const double *field,
                                const long long *field_off, const double *gobs, const long long *gobs_off,
                                const double *psi_hi, const long long *hi_off, const double *ckpt,
                                const long long *ckpt_off, int every, int n_ckpt, double *grad,
                                const long long *grad_off, void *stream)
{
    clear_status();
    if (n_inst < 0 || (n_inst > 0 && (!geom || !k_hi || !field || !field_off || !gobs || !gobs_off || !psi_hi ||
                                      !hi_off || !grad || !grad_off || (ckpt && !ckpt_off))))
        return set_status(CTDSE_ERR_ARGUMENT, "null adjoint argument");
    if (ckpt && (every < 1 || n_ckpt < 0)) return set_status(CTDSE_ERR_ARGUMENT, "bad checkpoint spacing");
    if (device_ready() != CTDSE_OK) return g_status;
    std::vector<const ctdse_geometry *> ig(n_inst);
    std::vector<TdseRevInst> ri(n_inst);
    for (int i = 0; i < n_inst; i++) {
        if (!geom[i] || k_hi[i] < 1)
            return set_status(CTDSE_ERR_ARGUMENT, "adjoint instance " + std::to_string(i) + ": bad range");
        ig[i] = geom[i];
        TdseRevInst &r = ri[i];
        r.k_hi = k_hi[i];
        r.field = field + field_off[i];
        r.gobs = gobs + gobs_off[i];
        r.psi_hi = (const double2 *)(psi_hi + hi_off[i]);
        r.ckpt = ckpt ? (const double2 *)(ckpt + ckpt_off[i]) : nullptr;
        r.every = ckpt ? every : 1;
        r.n_ckpt = ckpt ? n_ckpt : 0;
        r.grad = grad + grad_off[i];
    }
    return reverse_adjoint(ig, ri, (hipStream_t)stream);
}

void Initialise_grid_and_ground_state(inputs_def *in)
{
    clear_status();
    if (!in) {
        set_status(CTDSE_ERR_ARGUMENT, "null inputs");
        return;
    }
    if (in->num_r < 2 || !(in->dx > 0) || !(in->trg.a > 0)) {
        set_status(CTDSE_ERR_ARGUMENT, "num_r >= 2, dx > 0 and trg.a > 0 are required");
        return;
    }
    // one geometry's inverse iteration is a serial chain: it runs on the host, with the
    // reference's operation order (the same bits as tools.c:31-51 and the GPU lanes of
    // ctdse_geometry_create_batch)
    const size_t nx = in->num_r + 1;
    double *x = (double *)calloc(nx, sizeof(double));
    double *psi0 = (double *)calloc(2 * nx, sizeof(double));
    std::vector<double> work;
    try {
        work.resize(16 * (nx + 1));
    } catch (const std::bad_alloc &) {
        work.clear();
    }
    if (!x || !psi0 || work.empty()) {
        free(x);
        free(psi0);
        set_status(CTDSE_ERR_ARGUMENT, "Initialise_grid_and_ground_state: host allocation failed");
        return;
    }
    GsJob J{};
    J.num_r = in->num_r;
    J.max_iter = kGsMaxIter;
    J.dx = in->dx;
    J.a = in->trg.a;
    J.e_guess = in->Eguess;
    J.cv = in->CV;
    J.x = x;
    J.psi0 = psi0;
    J.work = work.data();
    tdse_ground_state_host(J);
    (void)gs_status(J, "");  // a limit cycle or a zero pivot is reported; the last iterate is kept
    in->x = x;
    in->psi0 = psi0;
    in->Einit = J.energy;
}

void call1DTDSE(inputs_def *inputs, outputs_def *outputs)
{
    clear_status();
    if (!inputs || !outputs) {
        set_status(CTDSE_ERR_ARGUMENT, "null inputs or outputs");
        return;
    }
    inputs_def &in = *inputs;
    if (!in.Efield.tgrid || !in.Efield.Field || in.Efield.Nt < 2) {
        set_status(CTDSE_ERR_ARGUMENT, "field and time grid with at least 2 samples are required");
        return;
    }
    if (!in.psi0 || !in.x) {
        set_status(CTDSE_ERR_ARGUMENT, "psi0 and x are required (call Initialise_grid_and_ground_state)");
        return;
    }
    if (in.gauge != 0) {
        set_status(CTDSE_ERR_UNSUPPORTED, "only the length gauge is supported");
        return;
    }
    if (!(in.dt > 0)) {
        set_status(CTDSE_ERR_ARGUMENT, "dt must be positive");
        return;
    }
    const int n_e = in.Efield.Nt;
    for (int i = 0; i < n_e; i++)
        if (!std::isfinite(in.Efield.Field[i])) {
            set_status(CTDSE_ERR_NONFINITE, "non-finite field sample " + std::to_string(i));
            return;
        }
    if (device_ready() != CTDSE_OK) return;

    // singleTDSE.c:52-68, 86-93
    const double field_dt = in.Efield.tgrid[1] - in.Efield.tgrid[0];
    const double tmax = in.Efield.tgrid[n_e - 1] - in.Efield.tgrid[0];
    int k1;
    if (in.InterpByDTorNT == 1) {
        k1 = in.Ntinterp + 1;
    } else {
        k1 = (int)std::floor(field_dt / in.dt);
        in.Ntinterp = k1;
        if (k1 != 1) k1++;
    }
    if (k1 < 1 || !(field_dt > 0)) {
        set_status(CTDSE_ERR_UNSUPPORTED, "field refinement factor k1 < 1 (field dt smaller than dt?)");
        return;
    }
    const double dt = field_dt / ((double)k1);
    const int n2 = k1 * n_e;
    const int nt_out = n2 + 2;
    in.Efield.dt = dt;
    in.num_t = (int)std::floor((2 * 3.14159265358979323846) / (0.057 * dt)) + 1;

    int n_snap = 0, spd = 1;
    if (in.analy.writewft) {
        spd = (int)std::floor(in.analy.tprint / dt);
        if (spd <= 0) {
            set_status(CTDSE_ERR_ARGUMENT, "tprint < dt with writewft set");
            return;
        }
        n_snap = nt_out / spd;
    }

    int st = CTDSE_OK;
    const std::shared_ptr<ctdse_geometry> geom = cached_geometry(in, dt, &st);
    if (!geom) return;
    ctdse_geometry *G = geom.get();
    const size_t nx = in.num_r + 1;
    DevMem fieldbuf, outbuf, snaps, tgrid;
    auto fail_dev = [&](hipError_t e) { set_status(CTDSE_ERR_DEVICE, std::string("call1DTDSE: ") + hipGetErrorString(e)); };
    hipError_t e = fieldbuf.alloc((size_t)(n2 + 3) * sizeof(double));
    if (e == hipSuccess) e = hipMemset(fieldbuf.p, 0, (size_t)(n2 + 3) * sizeof(double));
    if (e != hipSuccess) return fail_dev(e);
    if (k1 == 1) {
        // FourInterp with k = 1 is the identity up to FFT rounding (tools_fftw3.c:24-68)
        e = hipMemcpy(fieldbuf.p, in.Efield.Field, n_e * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail_dev(e);
    } else {
        DevMem raw, bins;
        const int nc = n_e / 2 + 1;
        e = raw.alloc(n_e * sizeof(double));
        if (e == hipSuccess) e = bins.alloc(nc * sizeof(double2));
        if (e == hipSuccess) e = hipMemcpy(raw.p, in.Efield.Field, n_e * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail_dev(e);
        hipLaunchKernelGGL(tdse_dft_r2c_kernel, dim3((nc + 127) / 128), dim3(128), 0, 0, raw.as<double>(), n_e, nc,
                           bins.as<double2>());
        hipLaunchKernelGGL(tdse_c2r_interp_kernel, dim3((n2 + 127) / 128), dim3(128), 0, 0, bins.as<double2>(), nc,
                           n_e, n2, fieldbuf.as<double>());
        if ((e = hipGetLastError()) != hipSuccess) return fail_dev(e);
    }
    e = outbuf.alloc(5 * (size_t)nt_out * sizeof(double));
    if (e == hipSuccess && n_snap > 0) e = snaps.alloc((size_t)n_snap * 2 * nx * sizeof(double));
    if (e == hipSuccess && n_snap > 0) e = hipMemset(snaps.p, 0, (size_t)n_snap * 2 * nx * sizeof(double));
    if (e == hipSuccess) e = tgrid.alloc((size_t)nt_out * sizeof(double));
    if (e != hipSuccess) return fail_dev(e);

    InstSpec s;
    s.geom = G;
    s.n_steps = nt_out;
    s.field = fieldbuf.as<double>();
    s.out = outbuf.as<double>();
    if (n_snap > 0) {
        s.snap = snaps.as<double>();
        s.snap_every = spd;
        s.n_snap = n_snap;
    }
    // one instance alone: its segments run side by side, one launch per step; where that
    // path cannot run, the persistent kernel with the widest segments (8 waves: the fewest
    // serial segments per step, 145 ms against 178 ms with 4 waves for 4416 steps of 16001
    // points)
    const ctdse_launch_config single{TDSE_MAX_WAVES, 1, 1};
    st = std::getenv("MMA_TDSE_NO_STEP_PARALLEL") ? (int)CTDSE_ERR_UNSUPPORTED : propagate_steps(s, 0);
    if (st == CTDSE_ERR_UNSUPPORTED) st = propagate({s}, 0, &single);
    if (st != CTDSE_OK) return;
    hipLaunchKernelGGL(tdse_tgrid_kernel, dim3(1), dim3(64), 0, 0, tgrid.as<double>(), nt_out, in.tmin, dt);

    // outputs: library-owned host arrays sized as singleTDSE.c:95-100 (Nt + 2 with Nt = n2 + 1)
    const size_t alloc_n = (size_t)n2 + 3;
    const double *o = outbuf.as<double>();
    outputs->tgrid = host_copy(tgrid.as<double>(), nt_out, alloc_n);
    outputs->sourceterm = host_copy(o, nt_out, alloc_n);
    outputs->PopTot = host_copy(o + nt_out, nt_out, alloc_n);
    outputs->PopInt = host_copy(o + 2 * (size_t)nt_out, nt_out, alloc_n);
    outputs->expval = host_copy(o + 3 * (size_t)nt_out, nt_out, alloc_n);
    outputs->Efield = host_copy(o + 4 * (size_t)nt_out, nt_out, alloc_n);
    outputs->Nt = nt_out;
    double *dummy = nullptr;
    st = spectrum_to_host(o + 4 * (size_t)nt_out, nt_out, dt, tmax, &dummy, &outputs->FEfield, &outputs->FEfieldM2);
    free(dummy);
    if (st == CTDSE_OK)
        st = spectrum_to_host(o, nt_out, dt, tmax, &outputs->omegagrid, &outputs->Fsourceterm,
                              &outputs->FsourcetermM2);
    outputs->Nomega = nt_out / 2 + 1;
    bool copied = outputs->tgrid && outputs->sourceterm && outputs->PopTot && outputs->PopInt && outputs->expval &&
                  outputs->Efield;
    if (in.analy.writewft) {
        outputs->psi = (double **)calloc(std::max(n_snap, 1), sizeof(double *));
        copied = copied && outputs->psi;
        for (int r = 0; outputs->psi && r < n_snap; r++) {
            outputs->psi[r] = host_copy(snaps.as<double>() + (size_t)r * 2 * nx, 2 * nx, 2 * nx);
            copied = copied && outputs->psi[r];
        }
    }
    // a failed copy leaves a NULL output: report it (the spectra set their own status)
    if (st == CTDSE_OK && !copied) set_status(CTDSE_ERR_DEVICE, "call1DTDSE: allocating or copying the outputs failed");
}

void set_time_and_field(inputs_def *in, double *time, double *field, int N)
{
    clear_status();
    if (!in || N < 0 || (N > 0 && (!time || !field))) {
        set_status(CTDSE_ERR_ARGUMENT, "set_time_and_field: bad arguments");
        return;
    }
    in->Efield.tgrid = (double *)calloc(std::max(N, 1), sizeof(double));
    in->Efield.Field = (double *)calloc(std::max(N, 1), sizeof(double));
    for (int i = 0; i < N; i++) {
        in->Efield.tgrid[i] = time[i];
        in->Efield.Field[i] = field[i];
    }
}

output_print_def Set_all_prints(void)
{
    output_print_def r;
    r.Efield = r.FEfield = r.sourceterm = r.Fsourceterm = r.FEfieldM2 = r.FsourceTermM2 = 1;
    r.PopTot = r.tgrid = r.omegagrid = r.PopInt = r.expval_x = 1;
    return r;
}

output_print_def Initialise_Printing_struct(void)
{
    output_print_def r;
    r.Efield = r.FEfield = r.sourceterm = r.Fsourceterm = r.FEfieldM2 = r.FsourceTermM2 = 0;
    r.PopTot = r.tgrid = r.omegagrid = r.PopInt = r.expval_x = 0;
    return r;
}

void free_arr(double *buffer) { free(buffer); }

void free_mtrx(double **buf, int N_rows)
{
    if (!buf) return;
    for (int i = 0; i < N_rows; i++) free(buf[i]);
    free(buf);
}

void outputs_destructor(outputs_def *o)
{
    if (!o) return;
    double **fields[] = {&o->tgrid, &o->omegagrid, &o->Efield, &o->sourceterm, &o->PopTot, &o->FEfield,
                         &o->Fsourceterm, &o->FEfieldM2, &o->FsourcetermM2, &o->PopInt, &o->expval};
    for (double **f : fields) {
        free(*f);
        *f = nullptr;
    }
}

void inputs_destructor(inputs_def *in)
{
    if (!in) return;
    free(in->psi0);
    in->psi0 = nullptr;
    free(in->x);
    in->x = nullptr;
    free(in->Efield.tgrid);
    in->Efield.tgrid = nullptr;
    free(in->Efield.Field);
    in->Efield.Field = nullptr;
}

} // extern "C"
