// lbm355_host.cpp - host side of liblbm355: case setup, brick index, C ABI.
//
// Replaces the host half of sim_tcpc (StateTcpc::setupBoundaries / resetLattice /
// updateKernelVelocities / statReset / outputData, solver/sim_tcpc/sim_tcpc.cu:
// 438-785) and of the lbm3d runtime (LBM container, execute, SimUpdate,
// solver/lbm3d/lbm.hpp, core.h:1029-1628).  The device work is in
// lbm355_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <sys/stat.h>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/lbm355.h"
#include "lbm355_internal.h"

using namespace lbm355;

namespace {

thread_local std::string g_error;

// a batch's graph capture and the device-memory management of case setup and
// teardown (hipMalloc, hipFree, synchronous copies) never overlap across
// threads: the serving pool creates and destroys cases on a builder thread
// while its worker captures graphs, and a device-wide synchronisation inside
// an allocator call must not meet a stream in capture
std::mutex g_device_mu;

int fail(int code, const std::string& msg)
{
    g_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) return fail(LBM355_EDEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCH_TRY(expr)                                                                \
    do {                                                                                \
        int r_ = (expr);                                                                \
        if (r_ != 0) return fail(LBM355_EDEVICE, std::string(#expr ": ") + kernel_error()); \
    } while (0)

inline int qx(int q) { return q / 9 - 1; }
inline int qy(int q) { return (q / 3) % 3 - 1; }
inline int qz(int q) { return q % 3 - 1; }

size_t storage_bytes(int s) { return s == LBM355_F64 ? 8 : s == LBM355_F32 ? 4 : 2; }
size_t compute_bytes(int s) { return s == LBM355_F64 ? 8 : 4; }

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t n)
    {
        bytes = n;
        if (n == 0) return 0;
        if (hipMalloc(&p, n) != hipSuccess) { p = nullptr; return -1; }
        return 0;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// the label setup of sim_tcpc: every cell fluid, x planes 3/4, y planes 5/6,
// z planes wall, wall mask from the geometry (value 1), open-boundary
// extrusion of the wall mask, areas, projectWall.  Generalised to a per-face
// table; all-zero faces reproduce sim_tcpc.cu:494-645 and lbm.hpp:145-155.
struct Labels {
    int X = 0, Y = 0, Z = 0;
    std::vector<uint8_t> map, wall;
    int64_t area[4] = {0, 0, 0, 0};  // x-, x+, y-, y+ open BC cells
    int64_t in_lab[4] = {0, 0, 0, 0};
    int64_t n_wall = 0, n_fluid = 0;
    int face[6];
    bool terminate = false;

    int64_t at(int x, int y, int z) const { return x + int64_t(X) * (y + int64_t(Y) * z); }

    int build(const lbm355_case_desc& d)
    {
        X = d.nx; Y = d.ny; Z = d.nz;
        const int64_t N = int64_t(X) * Y * Z;
        bool tcpc = true;
        for (int i = 0; i < 6; ++i) tcpc = tcpc && d.face_bc[i] == LBM355_FACE_TCPC;
        static const int layout[6] = {LBM355_FACE_INFLOW, LBM355_FACE_INFLOW, LBM355_FACE_OUTFLOW,
                                      LBM355_FACE_OUTFLOW, LBM355_FACE_WALL, LBM355_FACE_WALL};
        for (int i = 0; i < 6; ++i) {
            face[i] = tcpc ? layout[i] : d.face_bc[i];
            if (face[i] < LBM355_FACE_WALL || face[i] > LBM355_FACE_OUTFLOW)
                return fail(LBM355_EINVAL, "face_bc: mix of 0 and explicit kinds");
            if (face[i] == LBM355_FACE_INFLOW && i >= 2) return fail(LBM355_EINVAL, "inflow only on x faces");
            if (face[i] == LBM355_FACE_OUTFLOW && i >= 4) return fail(LBM355_EINVAL, "outflow only on x/y faces");
        }
        map.assign(N, L_FLUID);
        wall.resize(N);
        for (int64_t i = 0; i < N; ++i) {
            const int v = d.geometry[i];
            wall[i] = v == 1;
            if (v >= 3 && v <= 6) in_lab[v - 3]++;
        }
        const uint8_t xm = face[0] == LBM355_FACE_WALL ? L_WALL : face[0] == LBM355_FACE_INFLOW ? L_IN_XM : L_OUT_XM;
        const uint8_t xp = face[1] == LBM355_FACE_WALL ? L_WALL : face[1] == LBM355_FACE_INFLOW ? L_IN_XP : L_OUT_XP;
        const uint8_t ym = face[2] == LBM355_FACE_WALL ? L_WALL : L_OUT_YM;
        const uint8_t yp = face[3] == LBM355_FACE_WALL ? L_WALL : L_OUT_YP;
        for (int z = 0; z < Z; ++z)
            for (int y = 0; y < Y; ++y) { map[at(0, y, z)] = xm; map[at(X - 1, y, z)] = xp; }
        for (int z = 0; z < Z; ++z)
            for (int x = 0; x < X; ++x) { map[at(x, 0, z)] = ym; map[at(x, Y - 1, z)] = yp; }
        for (int y = 0; y < Y; ++y)
            for (int x = 0; x < X; ++x) { map[at(x, y, 0)] = L_WALL; map[at(x, y, Z - 1)] = L_WALL; }
        // open-boundary extrusion: copy the open pattern of the first non-empty
        // interior plane outwards to the boundary plane
        auto open_x = [&](int x) {
            for (int z = 0; z < Z; ++z)
                for (int y = 0; y < Y; ++y)
                    if (!wall[at(x, y, z)]) return true;
            return false;
        };
        auto open_y = [&](int y) {
            for (int z = 0; z < Z; ++z)
                for (int x = 0; x < X; ++x)
                    if (!wall[at(x, y, z)]) return true;
            return false;
        };
        auto extrude_x = [&](int bx, int dir) {
            int s = bx + dir;
            while (s > 0 && s < X - 1 && !open_x(s)) s += dir;
            if (s <= 0 || s >= X - 1) return;
            for (int z = 0; z < Z; ++z)
                for (int y = 0; y < Y; ++y)
                    if (!wall[at(s, y, z)])
                        for (int x = bx; x != s + dir; x += dir) wall[at(x, y, z)] = 0;
        };
        auto extrude_y = [&](int by, int dir) {
            int s = by + dir;
            while (s > 0 && s < Y - 1 && !open_y(s)) s += dir;
            if (s <= 0 || s >= Y - 1) return;
            for (int z = 0; z < Z; ++z)
                for (int x = 0; x < X; ++x)
                    if (!wall[at(x, s, z)])
                        for (int y = by; y != s + dir; y += dir) wall[at(x, y, z)] = 0;
        };
        if (face[0] != LBM355_FACE_WALL) extrude_x(0, 1);
        if (face[1] != LBM355_FACE_WALL) extrude_x(X - 1, -1);
        if (face[2] != LBM355_FACE_WALL) extrude_y(0, 1);
        if (face[3] != LBM355_FACE_WALL) extrude_y(Y - 1, -1);
        for (int64_t i = 0; i < N; ++i) {
            const int l = map[i];
            const bool w = wall[i] || l == L_WALL;
            if (!w) {
                if (l == L_IN_XM || l == L_OUT_XM) area[0]++;
                if (l == L_IN_XP || l == L_OUT_XP) area[1]++;
                if (l == L_OUT_YM) area[2]++;
                if (l == L_OUT_YP) area[3]++;
                if (l == L_FLUID) n_fluid++;
            } else {
                n_wall++;
            }
            if (wall[i]) map[i] = L_WALL;  // projectWall
        }
        for (int f = 0; f < 4; ++f)
            if (face[f] != LBM355_FACE_WALL && area[f] <= 0) terminate = true;
        return LBM355_OK;
    }
};

// ---------------------------------------------------------------------------
struct lbm355_case {
    lbm355_case_desc d{};
    Labels lab;
    int storage = 1, model = 0;
    int nbx = 0, nby = 0, nbz = 0;
    // brick origin shift: cell x sits in brick (x + ox) / kBX.  Chosen so that
    // the planes 0,1 and X-2,X-1 (and the same in y) share bricks: every BC
    // cell then lives in the wavefront of the cells whose slots it reads,
    // which lets the fused kernel read the BC partner populations itself.
    // z needs no shift: both z faces are always walls, so no BC cell has a
    // z-clamped partner (setup_device checks the property cell by cell).
    int ox = 0, oy = 0;
    std::vector<int32_t> brick_id;      // brick grid -> storage index or -1
    std::vector<int32_t> brick_xyz;     // storage index -> (bx,by,bz)
    int32_t n_active = 0, n_stored = 0, n_bc = 0;
    std::vector<int32_t> bc_cell;       // storage brick * 64 + lane of each BC slot
    std::vector<uint8_t> bc_label_h;
    double inflow[2] = {0, 0};
    int64_t steps = 0;
    int64_t macro_valid_step = -1;      // step index whose rho,u are in the macro buffer
    int64_t stats_steps = 0;            // samples accumulated since the last reset
    int64_t stats_first = 0;            // step index of the epoch (last reset or restart)
    int64_t stats_block0 = 0;           // stats_steps at the epoch: sample groups count from there
    bool strain_valid = false;

    hipStream_t stream = nullptr;
    int device = 0;                     // the HIP device the case lives on (current at creation)
    // recorded on the case's stream when a batch takes the case over, so that
    // the batch's steps start after everything already queued for the case
    hipEvent_t joined = nullptr;
    DevBuf f, stats, stats_part, macro, strain, force, nbr, label, bc_base, bc_cell_d, bc_read, bc_factor, bc_label,
        origin, base, devcase, list, flux;
    CaseDev host{};
    int bricks_per_group = 4;
    int nontemporal = -1;               // even-step nt hint: -1 auto (by working set), 0 off, 1 on
    int waves_cap = 0;                  // wavefronts per CU of the fused step, 0 = no cap

    ~lbm355_case()
    {
        if (stream) { (void)hipStreamSynchronize(stream); (void)hipStreamDestroy(stream); }
        if (joined) (void)hipEventDestroy(joined);
    }

    int64_t cell_index(int x, int y, int z) const { return x + int64_t(d.nx) * (y + int64_t(d.ny) * z); }
    int32_t stored_brick(int x, int y, int z) const
    {
        if (x < 0 || y < 0 || z < 0 || x >= d.nx || y >= d.ny || z >= d.nz) return -1;
        return brick_id[((x + ox) / kBX) + int64_t(nbx) * (((y + oy) / kBY) + int64_t(nby) * (z / kBZ))];
    }
    // element offset of slot `slot` of cell (x,y,z); the dummy brick if absent
    int lane_of(int x, int y, int z) const { return brick_lane(x + ox, y + oy, z); }
    void cell_of(int b, int l, int& x, int& y, int& z) const
    {
        x = brick_xyz[3 * b] * kBX + (l & (kBX - 1)) - ox;
        y = brick_xyz[3 * b + 1] * kBY + ((l >> kSX) & (kBY - 1)) - oy;
        z = brick_xyz[3 * b + 2] * kBZ + (l >> (kSX + kSY));
    }
    // the cell whose storage holds AA pair (cell, slot): the cell itself, or
    // for the two-byte types and slot < 13 the cell - c_slot (balanced
    // addressing, as the kernels' balanced_loads)
    void pair_position(int& x, int& y, int& z, int slot) const
    {
        if (balanced_storage(storage) && slot < 13) { x -= qx(slot); y -= qy(slot); z -= qz(slot); }
    }
    // element offset of AA pair (x,y,z; slot); the dummy brick if it has no storage
    int32_t slot_offset(int x, int y, int z, int slot) const
    {
        pair_position(x, y, z, slot);
        int32_t b = stored_brick(x, y, z);
        int lane = 0;
        if (b < 0) b = n_stored;
        else lane = lane_of(x, y, z);
        return (b * kQ + slot) * kLanes + lane;
    }
    // the same, or -1 if the pair has no storage (outside the grid or the
    // stored bricks): such a pair holds a population that no step reads
    int32_t pair_offset(int x, int y, int z, int slot) const
    {
        pair_position(x, y, z, slot);
        const int32_t b = stored_brick(x, y, z);
        return b < 0 ? -1 : (b * kQ + slot) * kLanes + lane_of(x, y, z);
    }
};

namespace {

// the partner reads of a BC cell: the reference pulls the populations of a
// virtual cell one plane inwards with the (clamped) indices below
// (sim_tcpc.cu:246,256,263,269; lbm_streaming.h:10-40).  For each direction q
// the source is cell C_q; under AA its post-collision f_q of the previous step
// sits in slot q of C_q + c_q after an odd step and in slot opp(q) of C_q
// after an even one.
void partner_sources(const lbm355_case& c, int x, int y, int z, int label, int (*src)[3])
{
    const int X = c.d.nx, Y = c.d.ny, Z = c.d.nz;
    const int xp = std::min(x + 1, X - 1), xm = std::max(x - 1, 0);
    const int yp = std::min(y + 1, Y - 1), ym = std::max(y - 1, 0);
    const int zp = std::min(z + 1, Z - 1), zm = std::max(z - 1, 0);
    int ax[3], ay[3], az[3];  // (xm', x', xp') passed to the streaming call
    ax[0] = xm; ax[1] = x; ax[2] = xp;
    ay[0] = ym; ay[1] = y; ay[2] = yp;
    az[0] = zm; az[1] = z; az[2] = zp;
    switch (label) {
    case L_IN_XM: case L_OUT_XM: ax[0] = x; ax[1] = xp; ax[2] = xp + 1; break;
    case L_IN_XP: case L_OUT_XP: ax[0] = xm - 1; ax[1] = xm; ax[2] = x; break;
    case L_OUT_YM: ay[0] = y; ay[1] = yp; ay[2] = yp + 1; break;
    case L_OUT_YP: ay[0] = ym - 1; ay[1] = ym; ay[2] = y; break;
    default: break;
    }
    for (int q = 0; q < kQ; ++q) {
        // f_q comes from the upstream index: cx=-1 -> xp', 0 -> x', +1 -> xm'
        src[q][0] = ax[1 - qx(q)];
        src[q][1] = ay[1 - qy(q)];
        src[q][2] = az[1 - qz(q)];
    }
}

int validate(const lbm355_case_desc* d)
{
    if (!d) return fail(LBM355_EINVAL, "null descriptor");
    if (d->nx < 4 || d->ny < 4 || d->nz < 4) return fail(LBM355_EINVAL, "grid must be at least 4 cells per axis");
    if (!d->geometry) return fail(LBM355_EINVAL, "null geometry");
    if (d->storage < 0 || d->storage > 3) return fail(LBM355_EINVAL, "unknown storage");
    if (d->collision < 0 || d->collision > 3) return fail(LBM355_EINVAL, "unknown collision");
    if (!(d->nu_ref > 0.0)) return fail(LBM355_EINVAL, "LBM viscosity must be positive (core.h:1044)");
    if (d->cy_enabled && (d->cy_model < 0 || d->cy_model > 2)) return fail(LBM355_EINVAL, "unknown cy_model");
    return LBM355_OK;
}

// device bytes of a case with the given brick counts
int64_t footprint_bytes(const lbm355_case_desc& d, int64_t n_active, int64_t n_stored, int64_t n_bc)
{
    const int64_t sb = storage_bytes(d.storage), cb = compute_bytes(d.storage);
    int64_t b = (n_stored + kSpareBricks) * kQ * kLanes * sb;   // populations
    b += (n_stored + kSpareBricks) * kLanes;                     // labels
    b += n_active * (kQ * 4 + 4 + 6);                            // nbr, bc_base, origin
    b += n_active * kMacroFields * kLanes * cb;                  // macro
    b += n_active * 6 * kLanes * cb;                             // strain
    if (d.cy_enabled && d.cy_model == LBM355_CY_FD_FORCE) b += n_active * 3 * kLanes * cb;  // body force
    if (d.stats_every > 0) b += n_active * kStatFields * kLanes * 8;
    if (d.stats_every > 0 && d.storage != LBM355_F64) b += (kStatBlock - 1) * n_active * kRingFields * kLanes * 4;
    b += n_bc * (4 * cb + 2 * kQ * 4 + 4 + 1 + 8);       // bc tables
    b += n_active * 8;                                    // brick list
    return b;
}

int build_index(lbm355_case& c)
{
    const lbm355_case_desc& d = c.d;
    const int X = d.nx, Y = d.ny, Z = d.nz;
    const Labels& L = c.lab;
    // planes 0,1 share a brick for any shift o < B - 1; planes n-2, n-1 do
    // unless n - 1 + o starts a brick.  Of the shifts that keep both pairs
    // together, take the one with the fewest active bricks (one wavefront
    // each): how the 8x8 tiles fall on the vessel cross-sections changes the
    // count by a few per cent (256x192x128 junction: 7,190 at (0,0), 7,066 at
    // the best shift).  Counted on runs of non-wall cells along x.
    auto valid_shift = [](int n, int B, int o) { return o < B - 1 && (n - 1 + o) % B != 0; };
    struct Run {
        int y, z, x0, x1;
    };
    std::vector<Run> runs;
    for (int z = 0; z < Z; ++z)
        for (int y = 0; y < Y; ++y)
            for (int x = 0; x < X;) {
                if (L.map[L.at(x, y, z)] == L_WALL) { ++x; continue; }
                const int x0 = x;
                while (x < X && L.map[L.at(x, y, z)] != L_WALL) ++x;
                runs.push_back(Run{y, z, x0, x});
            }
    const int mbx = (X + kBX - 1) / kBX + 1, mby = (Y + kBY - 1) / kBY + 1;
    const int mbz = (Z + kBZ - 1) / kBZ;
    std::vector<int32_t> stamp(size_t(mbx) * mby * mbz, -1);
    int64_t best = -1;
    c.ox = 0; c.oy = 0;
    for (int ox = 0; ox < kBX - 1; ++ox) {
        if (!valid_shift(X, kBX, ox)) continue;
        for (int oy = 0; oy < kBY - 1; ++oy) {
            if (!valid_shift(Y, kBY, oy)) continue;
            const int32_t tag = ox * kBY + oy;
            int64_t n = 0;
            for (const Run& r : runs) {
                const int64_t row = (int64_t(r.z / kBZ) * mby + (r.y + oy) / kBY) * mbx;
                for (int bx = (r.x0 + ox) / kBX; bx <= (r.x1 - 1 + ox) / kBX; ++bx) {
                    int32_t& t = stamp[size_t(row + bx)];
                    if (t != tag) { t = tag; ++n; }
                }
            }
            if (best < 0 || n < best) { best = n; c.ox = ox; c.oy = oy; }
        }
    }
    c.nbx = (X + c.ox + kBX - 1) / kBX; c.nby = (Y + c.oy + kBY - 1) / kBY; c.nbz = (Z + kBZ - 1) / kBZ;
    const int64_t NB = int64_t(c.nbx) * c.nby * c.nbz;
    // bricks touched by the AA update: the 3x3x3 neighbourhood of every
    // non-wall cell, plus the partner reads of BC cells
    std::vector<uint8_t> active_b(NB, 0), stored_b(NB, 0);
    auto bidx = [&](int x, int y, int z) {
        return ((x + c.ox) / kBX) + int64_t(c.nbx) * (((y + c.oy) / kBY) + int64_t(c.nby) * (z / kBZ));
    };
    for (int z = 0; z < Z; ++z)
        for (int y = 0; y < Y; ++y)
            for (int x = 0; x < X; ++x) {
                if (L.map[L.at(x, y, z)] == L_WALL) continue;
                active_b[bidx(x, y, z)] = 1;
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int a = x + dx, b = y + dy, e = z + dz;
                            if (a < 0 || b < 0 || e < 0 || a >= X || b >= Y || e >= Z) continue;
                            stored_b[bidx(a, b, e)] = 1;
                        }
                const int l = L.map[L.at(x, y, z)];
                if (is_bc_label(l)) {
                    int src[kQ][3];
                    partner_sources(c, x, y, z, l, src);
                    for (int q = 0; q < kQ; ++q) {
                        const int sx = src[q][0], sy = src[q][1], sz = src[q][2];
                        if (sx >= 0 && sy >= 0 && sz >= 0 && sx < X && sy < Y && sz < Z) stored_b[bidx(sx, sy, sz)] = 1;
                        const int tx = sx + qx(q), ty = sy + qy(q), tz = sz + qz(q);
                        if (tx >= 0 && ty >= 0 && tz >= 0 && tx < X && ty < Y && tz < Z) stored_b[bidx(tx, ty, tz)] = 1;
                    }
                }
            }
    // order: active bricks, then shell bricks, each z slab after z slab and
    // row-major (x fastest) inside a slab.  Only x and y shifts of the odd step
    // split population lines between bricks (a z shift moves whole lines), so
    // keeping each slab's bricks together keeps the wavefronts that share a
    // line close in launch order.  Measured against 3D Morton order: 19.4 vs
    // 20.0 us per step at 256x192x128, 152.7 vs 155.6 us at 512x384x256.
    std::vector<std::pair<uint64_t, int64_t>> act, shell;
    for (int bz = 0; bz < c.nbz; ++bz)
        for (int by = 0; by < c.nby; ++by)
            for (int bx = 0; bx < c.nbx; ++bx) {
                const int64_t i = bx + int64_t(c.nbx) * (by + int64_t(c.nby) * bz);
                if (active_b[i]) act.push_back({uint64_t(i), i});
                else if (stored_b[i]) shell.push_back({uint64_t(i), i});
            }
    c.n_active = int32_t(act.size());
    c.n_stored = int32_t(act.size() + shell.size());
    if (int64_t(c.n_stored + kSpareBricks) * kQ * kLanes * int64_t(storage_bytes(d.storage)) >= (int64_t(1) << 32))
        return fail(LBM355_EINVAL, "case too large: populations must stay below 4 GiB per case");
    c.brick_id.assign(NB, -1);
    c.brick_xyz.resize(size_t(c.n_stored) * 3);
    int32_t n = 0;
    for (auto* v : {&act, &shell})
        for (auto& kv : *v) {
            const int64_t i = kv.second;
            c.brick_id[i] = n;
            c.brick_xyz[3 * n + 0] = int32_t(i % c.nbx);
            c.brick_xyz[3 * n + 1] = int32_t((i / c.nbx) % c.nby);
            c.brick_xyz[3 * n + 2] = int32_t(i / (int64_t(c.nbx) * c.nby));
            ++n;
        }
    return LBM355_OK;
}

template <class T>
int upload(DevBuf& buf, const std::vector<T>& v)
{
    if (buf.alloc(std::max<size_t>(v.size() * sizeof(T), 16)) != 0) return fail(LBM355_ENOMEM, "hipMalloc failed");
    if (!v.empty()) HIP_TRY(hipMemcpy(buf.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return LBM355_OK;
}

int alloc_zero(lbm355_case& c, DevBuf& buf, size_t bytes)
{
    if (buf.alloc(std::max<size_t>(bytes, 16)) != 0) return fail(LBM355_ENOMEM, "hipMalloc failed");
    LAUNCH_TRY(launch_zero(buf.p, int64_t((std::max<size_t>(bytes, 16) + 3) / 4 * 4), c.stream));
    return LBM355_OK;
}

int setup_device(lbm355_case& c)
{
    const lbm355_case_desc& d = c.d;
    const Labels& L = c.lab;
    const int X = d.nx, Y = d.ny, Z = d.nz;
    const int na = c.n_active, ns = c.n_stored;
    HIP_TRY(hipGetDevice(&c.device));
    HIP_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&c.joined, hipEventDisableTiming));

    // labels per stored brick (+ dummy and sink = wall)
    std::vector<uint8_t> lab(size_t(ns + kSpareBricks) * kLanes, L_WALL);
    for (int b = 0; b < ns; ++b)
        for (int l = 0; l < kLanes; ++l) {
            int x, y, z;
            c.cell_of(b, l, x, y, z);
            lab[size_t(b) * kLanes + l] =
                (x >= 0 && y >= 0 && x < X && y < Y && z < Z) ? L.map[L.at(x, y, z)] : L_OUTSIDE;
        }
    // neighbour bricks of the active bricks; absent ones are the sink, which
    // takes the odd step's writes that leave the grid
    const int32_t sink = (ns + 1) * kQ * kLanes;
    std::vector<int32_t> nbr(size_t(std::max(na, 1)) * kQ, sink);
    std::vector<int16_t> org(size_t(std::max(na, 1)) * 3, 0);
    for (int b = 0; b < na; ++b) {
        const int bx = c.brick_xyz[3 * b], by = c.brick_xyz[3 * b + 1], bz = c.brick_xyz[3 * b + 2];
        org[3 * b] = int16_t(bx); org[3 * b + 1] = int16_t(by); org[3 * b + 2] = int16_t(bz);
        for (int o = 0; o < kQ; ++o) {
            const int nx = bx + qx(o), ny = by + qy(o), nz = bz + qz(o);
            int32_t s = -1;
            if (nx >= 0 && ny >= 0 && nz >= 0 && nx < c.nbx && ny < c.nby && nz < c.nbz)
                s = c.brick_id[nx + int64_t(c.nbx) * (ny + int64_t(c.nby) * nz)];
            // element offset of the neighbour brick's population block
            nbr[size_t(b) * kQ + o] = s < 0 ? sink : s * kQ * kLanes;
        }
    }
    // BC slots in (brick, lane) order
    std::vector<int32_t> bc_base(size_t(std::max(na, 1)), -1), bc_read;
    std::vector<double> bc_factor;
    c.bc_cell.clear();
    c.bc_label_h.clear();
    // inflow profile: centre and radius per x face
    double cen[2][2], rad[2];
    for (int side = 0; side < 2; ++side) {
        const int x = side == 0 ? 0 : X - 1;
        const uint8_t want = side == 0 ? L_IN_XM : L_IN_XP;
        const double* cc = side == 0 ? d.ivc_center : d.svc_center;
        double sy = 0, sz = 0;
        int64_t cnt = 0;
        for (int z = 0; z < Z; ++z)
            for (int y = 0; y < Y; ++y)
                if (L.map[L.at(x, y, z)] == want) { sy += y; sz += z; ++cnt; }
        cen[side][0] = cc[0] < 0 && cnt ? sy / double(cnt) : cc[0];
        cen[side][1] = cc[1] < 0 && cnt ? sz / double(cnt) : cc[1];
        double R = side == 0 ? d.ivc_radius : d.svc_radius;
        if (R <= 0) {
            double r2 = 0;
            for (int z = 0; z < Z; ++z)
                for (int y = 0; y < Y; ++y)
                    if (L.map[L.at(x, y, z)] == want)
                        r2 = std::max(r2, (y - cen[side][0]) * (y - cen[side][0]) + (z - cen[side][1]) * (z - cen[side][1]));
            R = std::sqrt(r2) + 0.5;
        }
        rad[side] = R;
    }
    for (int b = 0; b < na; ++b) {
        bc_base[b] = int32_t(c.bc_cell.size());
        for (int l = 0; l < kLanes; ++l) {
            const int lb = lab[size_t(b) * kLanes + l];
            if (!is_bc_label(lb)) continue;
            int x, y, z;
            c.cell_of(b, l, x, y, z);
            c.bc_cell.push_back(b * kLanes + l);
            c.bc_label_h.push_back(uint8_t(lb));
            int src[kQ][3];
            partner_sources(c, x, y, z, lb, src);
            // the slot read for q is rewritten in the same step by cell
            // C_q + c_q (either parity); that cell must be a lane of this
            // wavefront so that the read precedes the write.  The brick origin
            // shift guarantees it on x and y faces, and z faces are walls.
            for (int q = 0; q < kQ; ++q) {
                const int wx = src[q][0] + qx(q), wy = src[q][1] + qy(q), wz = src[q][2] + qz(q);
                if (wx < 0 || wy < 0 || wz < 0 || wx >= X || wy >= Y || wz >= Z) continue;
                if (L.map[L.at(wx, wy, wz)] != L_WALL && c.stored_brick(wx, wy, wz) != b)
                    return fail(LBM355_EINVAL, "boundary cell (" + std::to_string(x) + "," + std::to_string(y) + "," +
                                                   std::to_string(z) +
                                                   "): a partner slot is written outside its wavefront");
            }
            // parity 0 (this step even, the previous one odd): slot q of C_q + c_q
            for (int q = 0; q < kQ; ++q)
                bc_read.push_back(c.slot_offset(src[q][0] + qx(q), src[q][1] + qy(q), src[q][2] + qz(q), q));
            // parity 1: slot opp(q) of C_q
            for (int q = 0; q < kQ; ++q) bc_read.push_back(c.slot_offset(src[q][0], src[q][1], src[q][2], 26 - q));
            double fac = 1.0;
            if (is_inflow_label(lb)) {
                const int side = lb == L_IN_XM ? 0 : 1;
                if (d.profile == LBM355_PROFILE_PARABOLIC) {
                    const double r2 = ((y - cen[side][0]) * (y - cen[side][0]) + (z - cen[side][1]) * (z - cen[side][1])) /
                                      (rad[side] * rad[side]);
                    fac = std::max(0.0, 1.0 - r2);
                } else {
                    const int64_t area = L.area[side];
                    fac = area > 0 ? double(area) : 1.0e300;  // plug: u = flux / area
                }
            }
            bc_factor.push_back(fac);
        }
    }
    c.n_bc = int32_t(c.bc_cell.size());

    const size_t sb = storage_bytes(d.storage), cb = compute_bytes(d.storage);
    if (c.f.alloc(size_t(ns + kSpareBricks) * kQ * kLanes * sb) != 0) return fail(LBM355_ENOMEM, "population buffer");
    LAUNCH_TRY(launch_fill_populations(d.storage, c.f.p, ns + kSpareBricks, c.stream));
    int rc;
    if ((rc = upload(c.label, lab))) return rc;
    if ((rc = upload(c.nbr, nbr))) return rc;
    if ((rc = upload(c.origin, org))) return rc;
    if ((rc = upload(c.bc_base, bc_base))) return rc;
    if ((rc = upload(c.bc_read, bc_read))) return rc;
    if ((rc = upload(c.bc_factor, bc_factor))) return rc;
    if ((rc = upload(c.bc_label, c.bc_label_h))) return rc;
    if ((rc = upload(c.bc_cell_d, c.bc_cell))) return rc;
    // rho = 1, u = 0 for the initial macro state
    if ((rc = alloc_zero(c, c.macro, size_t(std::max(na, 1)) * kMacroFields * kLanes * cb))) return rc;
    if ((rc = alloc_zero(c, c.strain, size_t(std::max(na, 1)) * 6 * kLanes * cb))) return rc;
    const bool fd_force = d.cy_enabled && d.cy_model == LBM355_CY_FD_FORCE;
    if (fd_force && (rc = alloc_zero(c, c.force, size_t(std::max(na, 1)) * 3 * kLanes * cb))) return rc;
    if (d.stats_every > 0 && (rc = alloc_zero(c, c.stats, size_t(std::max(na, 1)) * kStatFields * kLanes * 8))) return rc;
    // ring of FP32 rho u samples, one slot per sample of the open group
    // (FP64 storage adds every sample to the FP64 sums directly)
    const bool ring = d.stats_every > 0 && d.storage != LBM355_F64;
    if (ring && (rc = alloc_zero(c, c.stats_part, size_t(kStatBlock - 1) * std::max(na, 1) * kRingFields * kLanes * 4)))
        return rc;
    if ((rc = alloc_zero(c, c.base, 8))) return rc;
    if ((rc = alloc_zero(c, c.flux, 4 * 8))) return rc;
    std::vector<BrickRef> list(size_t(std::max(na, 1)));
    for (int b = 0; b < na; ++b) list[b] = BrickRef{0u, uint32_t(b)};
    if ((rc = upload(c.list, list))) return rc;

    CaseDev& h = c.host;
    std::memset(&h, 0, sizeof(h));
    h.f = c.f.p;
    h.stats = static_cast<double*>(d.stats_every > 0 ? c.stats.p : nullptr);
    h.stats_part = static_cast<float*>(ring ? c.stats_part.p : nullptr);
    h.stats_epoch = 0;
    h.macro = c.macro.p;
    h.strain = c.strain.p;
    h.force = fd_force ? c.force.p : nullptr;
    h.nbr = static_cast<const int32_t*>(c.nbr.p);
    h.label = static_cast<const uint8_t*>(c.label.p);
    h.bc_base = static_cast<const int32_t*>(c.bc_base.p);
    h.bc_cell = static_cast<const int32_t*>(c.bc_cell_d.p);
    h.bc_read = static_cast<const int32_t*>(c.bc_read.p);
    h.bc_factor = static_cast<const double*>(c.bc_factor.p);
    h.plug = d.profile == LBM355_PROFILE_PARABOLIC ? 0 : 1;
    h.bc_label = static_cast<const uint8_t*>(c.bc_label.p);
    h.origin = static_cast<const int16_t*>(c.origin.p);
    h.start = 0;
    h.f_bytes = int64_t(c.f.bytes);
    h.stats_bytes = int64_t(c.stats.bytes);
    h.n_active = na;
    h.n_stored = ns;
    h.n_bc = c.n_bc;
    h.nx = X; h.ny = Y; h.nz = Z;
    h.nu_ref = d.nu_ref;
    h.cy_nu0 = d.cy_nu0; h.cy_lambda = d.cy_lambda; h.cy_a = d.cy_a; h.cy_n = d.cy_n;
    {
        const double tr = 3.0 * d.nu_ref + 0.5, t0 = 3.0 * d.cy_nu0 + 0.5;
        h.cy_e = (d.cy_n - 1.0) / d.cy_a;
        h.cy_log2_lambda = std::log2(d.cy_lambda);
        h.cy_log2_tau_max = std::log2(std::max(tr, t0));
        h.cy_log2_tau_min = std::log2(std::min(tr, t0));
        h.nu_ref_f = float(d.nu_ref);
        h.cy_nu0_f = float(d.cy_nu0); h.cy_a_f = float(d.cy_a); h.cy_n_f = float(d.cy_n);
        h.cy_e_f = float(h.cy_e);
        h.cy_log2_lambda_f = float(h.cy_log2_lambda);
        h.cy_log2_tau_max_f = float(h.cy_log2_tau_max);
        h.cy_log2_tau_min_f = float(h.cy_log2_tau_min);
    }
    h.cy_mode = d.cy_enabled ? d.cy_model : 0;
    h.inflow_xm = c.inflow[0];
    h.inflow_xp = c.inflow[1];
    h.pulse_amp = d.pulse_amp;
    h.pulse_period = d.pulse_period;
    h.stats_every = d.stats_every;
    h.macro_step = -1;
    h.macro_every = 0;
    if (c.devcase.alloc(sizeof(CaseDev)) != 0) return fail(LBM355_ENOMEM, "case constants");
    HIP_TRY(hipMemcpy(c.devcase.p, &h, sizeof(CaseDev), hipMemcpyHostToDevice));
    HIP_TRY(hipStreamSynchronize(c.stream));
    return LBM355_OK;
}

// device field of the case struct, for single-value updates
void* case_field(lbm355_case& c, const void* host_field)
{
    const size_t off = static_cast<const char*>(host_field) - reinterpret_cast<const char*>(&c.host);
    return static_cast<char*>(c.devcase.p) + off;
}

int set_case_u64(lbm355_case& c, void* host_field, uint64_t v)
{
    std::memcpy(host_field, &v, 8);
    LAUNCH_TRY(launch_set_u64(case_field(c, host_field), v, c.stream));
    return LBM355_OK;
}

int set_case_f64(lbm355_case& c, double* host_field, double v)
{
    uint64_t u;
    std::memcpy(&u, &v, 8);
    return set_case_u64(c, host_field, u);
}

StepLaunch single_launch(lbm355_case& c);

// device memory held by a case
int64_t case_device_bytes(const lbm355_case& c)
{
    int64_t bytes = 0;
    for (const DevBuf* b : {&c.f, &c.stats, &c.stats_part, &c.macro, &c.strain, &c.force, &c.nbr, &c.label, &c.bc_base,
                            &c.bc_cell_d, &c.bc_read, &c.bc_factor, &c.bc_label, &c.origin, &c.base,
                            &c.devcase, &c.list, &c.flux})
        bytes += int64_t(b->bytes);
    return bytes;
}

// device memory a step of the case sweeps: all of it except the statistics
// ring, of which a step touches one slot (the group's flush reads the rest
// once every kStatBlock samples).  Cache-residency decisions (the automatic
// non-temporal even step, batch groups, the pool's concurrent ticks) use this;
// admission against a memory budget uses case_device_bytes.
int64_t case_step_bytes(const lbm355_case& c)
{
    const int64_t ring = int64_t(c.stats_part.bytes);
    return case_device_bytes(c) - ring + ring / (kStatBlock - 1);
}

// one time step of a case: the fd-force pre-kernels (velocity, strain, body
// force) when that Carreau-Yasuda mode is on, then the fused step
int launch_case_step(lbm355_case& c, const StepLaunch& L)
{
    if (c.host.force) {
        const int r = launch_fd_force_prepass(c.storage, L, static_cast<const CaseDev*>(c.devcase.p), c.n_active,
                                              c.stream);
        if (r) return r;
    }
    return launch_step(c.storage, c.model, L, c.stream);
}

// dynamic LDS per workgroup that leaves room for at most cap / bpg
// workgroups (cap wavefronts) per CU; 0 for no cap.  The fused step uses no
// LDS itself, the allocation only limits how many of its wavefronts share a
// CU and hence the L2 footprint of the odd step's neighbour accesses.
int occupancy_lds(int cap, int bpg)
{
    if (cap <= 0) return 0;
    struct Lds {
        int per_cu, per_block;
    };
    // queried once (thread-safe static initialisation); every GPU this
    // library drives is the same model
    static const Lds lds = [] {
        int dev = 0, v = 0;
        Lds r{160 * 1024, 64 * 1024};
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess && v > 0)
            r.per_cu = v;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            r.per_block = v;
        return r;
    }();
    const int groups = std::max(1, cap / bpg);
    return std::min(lds.per_cu / (groups + 1) + 1, lds.per_block);
}

StepLaunch single_launch(lbm355_case& c)
{
    StepLaunch L{};
    L.list = static_cast<const BrickRef*>(c.list.p);
    L.n_list = c.n_active;
    L.cases = static_cast<CaseDev*>(c.devcase.p);
    L.n_cases = 1;
    L.base = static_cast<const int64_t*>(c.base.p);
    L.rel = c.steps;
    L.bricks_per_group = c.bricks_per_group;
    L.nontemporal = c.nontemporal >= 0 ? c.nontemporal : (case_step_bytes(c) > kNontemporalBytes ? 1 : 0);
    L.lds_bytes = occupancy_lds(c.waves_cap, c.bricks_per_group);
    return L;
}

// slot sl of the statistics ring: [active brick][3][64] FP32 rho u samples
const float* ring_slot(const lbm355_case& c, int64_t sl)
{
    return static_cast<const float*>(c.stats_part.p) + sl * int64_t(std::max(c.n_active, 1)) * kRingFields * kLanes;
}

// copy a per-active-brick device array [brick][nf][64] of type R to dense
// double arrays; `fill` for cells without an active brick
template <class R>
int gather_dense(lbm355_case& c, const void* dev, int nf, int field, double fill, double* out)
{
    const int na = c.n_active;
    std::vector<R> h(size_t(std::max(na, 1)) * nf * kLanes);
    HIP_TRY(hipStreamSynchronize(c.stream));
    if (na > 0) HIP_TRY(hipMemcpy(h.data(), dev, h.size() * sizeof(R), hipMemcpyDeviceToHost));
    const int X = c.d.nx, Y = c.d.ny, Z = c.d.nz;
    const int64_t N = int64_t(X) * Y * Z;
    std::fill(out, out + N, fill);
    for (int b = 0; b < na; ++b)
        for (int l = 0; l < kLanes; ++l) {
            int x, y, z;
            c.cell_of(b, l, x, y, z);
            if (x < 0 || y < 0 || x >= X || y >= Y || z >= Z) continue;
            const int64_t i = c.cell_index(x, y, z);
            if (c.lab.map[i] == L_WALL) continue;
            out[i] = double(h[(size_t(b) * nf + field) * kLanes + l]);
        }
    return LBM355_OK;
}

// inverse of gather_dense: overwrite the entries of the active cells of one
// field of a per-active-brick device array from a dense double array
template <class R>
int scatter_dense(lbm355_case& c, void* dev, int nf, int field, const double* in)
{
    const int na = c.n_active;
    if (na <= 0) return LBM355_OK;
    std::vector<R> h(size_t(na) * nf * kLanes);
    HIP_TRY(hipStreamSynchronize(c.stream));
    HIP_TRY(hipMemcpy(h.data(), dev, h.size() * sizeof(R), hipMemcpyDeviceToHost));
    const int X = c.d.nx, Y = c.d.ny, Z = c.d.nz;
    for (int b = 0; b < na; ++b)
        for (int l = 0; l < kLanes; ++l) {
            int x, y, z;
            c.cell_of(b, l, x, y, z);
            if (x < 0 || y < 0 || x >= X || y >= Y || z >= Z) continue;
            const int64_t i = c.cell_index(x, y, z);
            if (c.lab.map[i] == L_WALL) continue;
            h[(size_t(b) * nf + field) * kLanes + l] = R(in[i]);
        }
    HIP_TRY(hipMemcpy(dev, h.data(), h.size() * sizeof(R), hipMemcpyHostToDevice));
    return LBM355_OK;
}

int refresh_macro(lbm355_case& c)
{
    if (c.macro_valid_step == c.steps - 1 || c.steps == 0) return LBM355_OK;
    return fail(LBM355_EINVAL, "macroscopic fields of the last step were not written");
}

// big-endian writers of the legacy VTK format (vtk_writer.hpp:1-60)
void put_be32(std::vector<char>& buf, uint32_t v)
{
    buf.push_back(char(v >> 24)); buf.push_back(char(v >> 16)); buf.push_back(char(v >> 8)); buf.push_back(char(v));
}
void put_float(std::vector<char>& buf, float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    put_be32(buf, u);
}


// fn(offset, q, dense index q * N + cell) for every population the AA buffer
// holds for the case's current step parity: the post-collision f_q of a cell
// is pair (x + c_q, q) after an odd step and pair (x, opp q) after an even
// one; wall cells next to the flow hold the populations in flight through
// full-way bounce-back, the other wall cells hold nothing
template <class Fn>
void for_each_held(const lbm355_case& c, Fn&& fn)
{
    const int X = c.d.nx, Y = c.d.ny, Z = c.d.nz;
    const int64_t N = int64_t(X) * Y * Z;
    const bool even_next = (c.steps & 1) == 0;  // the next step is even: the last one was odd
    for (int z = 0; z < Z; ++z)
        for (int y = 0; y < Y; ++y)
            for (int x = 0; x < X; ++x) {
                const int64_t i = c.cell_index(x, y, z);
                if (c.lab.map[i] == L_WALL && c.stored_brick(x, y, z) < 0) continue;
                for (int q = 0; q < kQ; ++q) {
                    const int32_t off = even_next ? c.pair_offset(x + qx(q), y + qy(q), z + qz(q), q)
                                                  : c.pair_offset(x, y, z, 26 - q);
                    if (off >= 0) fn(off, q, int64_t(q) * N + i);
                }
            }
}
}  // namespace

// ===========================================================================
extern "C" {

const char* lbm355_version(void) { return "lbm355 1.0 (gfx950, AA bricks, ABI 1)"; }
const char* lbm355_last_error(void) { return g_error.c_str(); }

int lbm355_device_count(int* count)
{
    if (!count) return fail(LBM355_EINVAL, "null count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return LBM355_OK;
}

int lbm355_case_footprint(const lbm355_case_desc* desc, int64_t* bytes)
{
    int rc = validate(desc);
    if (rc) return rc;
    if (!bytes) return fail(LBM355_EINVAL, "null bytes");
    lbm355_case c;
    c.d = *desc;
    if ((rc = c.lab.build(c.d))) return rc;
    if ((rc = build_index(c))) return rc;
    int64_t nbc = 0;
    for (int i = 0; i < 4; ++i) nbc += c.lab.area[i];
    *bytes = footprint_bytes(c.d, c.n_active, c.n_stored, nbc);
    return LBM355_OK;
}

int lbm355_case_labels(const lbm355_case_desc* desc, int32_t* labels, lbm355_case_info* o)
{
    int rc = validate(desc);
    if (rc) return rc;
    lbm355_case c;
    c.d = *desc;
    if ((rc = c.lab.build(c.d))) return rc;
    if ((rc = build_index(c))) return rc;
    const int64_t N = int64_t(desc->nx) * desc->ny * desc->nz;
    if (labels)
        for (int64_t i = 0; i < N; ++i) labels[i] = c.lab.map[i];
    if (o) {
        std::memset(o, 0, sizeof(*o));
        o->fluid = c.lab.n_fluid;
        o->wall = c.lab.n_wall;
        o->area_ivc = c.lab.area[0];
        o->area_svc = c.lab.area[1];
        o->area_out5 = c.lab.area[2];
        o->area_out6 = c.lab.area[3];
        o->input_label3 = c.lab.in_lab[0];
        o->input_label4 = c.lab.in_lab[1];
        o->input_label5 = c.lab.in_lab[2];
        o->input_label6 = c.lab.in_lab[3];
        o->active_cells = N - c.lab.n_wall;
        o->bricks_active = c.n_active;
        o->bricks_stored = c.n_stored;
        int64_t nbc = 0;
        for (int i = 0; i < 4; ++i) nbc += c.lab.area[i];
        o->device_bytes = footprint_bytes(c.d, c.n_active, c.n_stored, nbc);
        o->terminate = c.lab.terminate ? 1 : 0;
    }
    return c.lab.terminate ? fail(LBM355_ETERMINATE, "zero inlet or outlet area") : LBM355_OK;
}

int lbm355_case_create(const lbm355_case_desc* desc, lbm355_case** out)
{
    int rc = validate(desc);
    if (rc) return rc;
    if (!out) return fail(LBM355_EINVAL, "null out");
    *out = nullptr;
    std::unique_ptr<lbm355_case> c(new (std::nothrow) lbm355_case());
    if (!c) return fail(LBM355_ENOMEM, "host allocation");
    c->d = *desc;
    c->d.geometry = nullptr;
    c->storage = desc->storage;
    c->model = desc->collision;
    c->inflow[0] = desc->profile == LBM355_PROFILE_PARABOLIC ? desc->ivc_umax : desc->ivc_flux;
    c->inflow[1] = desc->profile == LBM355_PROFILE_PARABOLIC ? desc->svc_umax : desc->svc_flux;
    {
        lbm355_case_desc tmp = *desc;
        if ((rc = c->lab.build(tmp))) return rc;
    }
    if ((rc = build_index(*c))) return rc;
    {
        std::lock_guard<std::mutex> device_lock(g_device_mu);
        if ((rc = setup_device(*c))) return rc;
    }
    const bool term = c->lab.terminate;
    *out = c.release();
    if (term) return fail(LBM355_ETERMINATE, "zero inlet or outlet area");
    return LBM355_OK;
}

int lbm355_case_destroy(lbm355_case* c)
{
    std::lock_guard<std::mutex> device_lock(g_device_mu);
    delete c;
    return LBM355_OK;
}

int lbm355_case_info_get(const lbm355_case* c, lbm355_case_info* o)
{
    if (!c || !o) return fail(LBM355_EINVAL, "null argument");
    std::memset(o, 0, sizeof(*o));
    o->fluid = c->lab.n_fluid;
    o->wall = c->lab.n_wall;
    o->area_ivc = c->lab.area[0];
    o->area_svc = c->lab.area[1];
    o->area_out5 = c->lab.area[2];
    o->area_out6 = c->lab.area[3];
    o->input_label3 = c->lab.in_lab[0];
    o->input_label4 = c->lab.in_lab[1];
    o->input_label5 = c->lab.in_lab[2];
    o->input_label6 = c->lab.in_lab[3];
    o->active_cells = int64_t(c->d.nx) * c->d.ny * c->d.nz - c->lab.n_wall;
    o->bricks_active = c->n_active;
    o->bricks_stored = c->n_stored;
    o->device_bytes = case_device_bytes(*c);
    o->steps = c->steps;
    o->terminate = c->lab.terminate ? 1 : 0;
    return LBM355_OK;
}

int lbm355_case_set_inflow(lbm355_case* c, double ivc, double svc)
{
    if (!c) return fail(LBM355_EINVAL, "null case");
    c->inflow[0] = ivc;
    c->inflow[1] = svc;
    int rc;
    if (c->host.inflow_xm != ivc && (rc = set_case_f64(*c, &c->host.inflow_xm, ivc))) return rc;
    if (c->host.inflow_xp != svc && (rc = set_case_f64(*c, &c->host.inflow_xp, svc))) return rc;
    return LBM355_OK;
}

int lbm355_case_step(lbm355_case* c, int64_t n)
{
    if (!c) return fail(LBM355_EINVAL, "null case");
    if (n <= 0) return LBM355_OK;
    int rc;
    // rho,u of the last step of this call are kept for get_field / VTK
    if ((rc = set_case_u64(*c, &c->host.macro_step, uint64_t(c->steps + n - 1)))) return rc;
    StepLaunch L = single_launch(*c);
    for (int64_t s = 0; s < n; ++s) {
        L.rel = c->steps + s;
        LAUNCH_TRY(launch_case_step(*c, L));
    }
    if (c->d.stats_every > 0) {
        const int64_t k = c->d.stats_every;
        c->stats_steps += (c->steps + n) / k - c->steps / k;
    }
    c->steps += n;
    c->macro_valid_step = c->steps - 1;
    c->strain_valid = false;
    return LBM355_OK;
}

int lbm355_case_set_shape(lbm355_case* c, int bpg)
{
    if (!c || !(bpg == 1 || bpg == 2 || bpg == 4 || bpg == 8)) return fail(LBM355_EINVAL, "bricks per group");
    c->bricks_per_group = bpg;
    return LBM355_OK;
}

int lbm355_case_set_nontemporal(lbm355_case* c, int mode)
{
    if (!c || mode < -1 || mode > 1) return fail(LBM355_EINVAL, "nontemporal mode is -1, 0 or 1");
    c->nontemporal = mode;
    return LBM355_OK;
}

int lbm355_case_set_occupancy(lbm355_case* c, int waves_per_cu)
{
    if (!c || !(waves_per_cu == 0 || (waves_per_cu >= 4 && waves_per_cu <= 32)))
        return fail(LBM355_EINVAL, "wavefronts per CU is 0 (no cap) or 4..32");
    c->waves_cap = waves_per_cu;
    return LBM355_OK;
}

int lbm355_case_sync(lbm355_case* c)
{
    if (!c) return fail(LBM355_EINVAL, "null case");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM355_OK;
}

// HIP events and a stream released on every return path
struct EventSet {
    std::vector<hipEvent_t> ev;
    explicit EventSet(size_t n) : ev(n, nullptr) {}
    ~EventSet()
    {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};
struct StreamHolder {
    hipStream_t s = nullptr;
    ~StreamHolder()
    {
        if (s) (void)hipStreamDestroy(s);
    }
};

int lbm355_case_time_kernel(lbm355_case* c, int64_t n, double* avg_step_ms, double* avg_kernel_ms)
{
    if (!c || n <= 0) return fail(LBM355_EINVAL, "bad argument");
    int rc;
    const int64_t n1 = (n + 1) / 2, n2 = n - n1;  // back to back, then per-launch events
    if ((rc = set_case_u64(*c, &c->host.macro_step, uint64_t(c->steps + n - 1)))) return rc;
    StepLaunch L = single_launch(*c);
    const int64_t nk = std::max<int64_t>(1, std::min<int64_t>(n2, 512));
    EventSet es(size_t(2 * nk + 2));
    std::vector<hipEvent_t>& ev = es.ev;
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventRecord(ev[2 * nk], c->stream));
    for (int64_t s = 0; s < n1; ++s) {
        L.rel = c->steps + s;
        LAUNCH_TRY(launch_case_step(*c, L));
    }
    HIP_TRY(hipEventRecord(ev[2 * nk + 1], c->stream));
    for (int64_t s = 0; s < n2; ++s) {
        L.rel = c->steps + n1 + s;
        if (s < nk) HIP_TRY(hipEventRecord(ev[2 * s], c->stream));
        LAUNCH_TRY(launch_case_step(*c, L));
        if (s < nk) HIP_TRY(hipEventRecord(ev[2 * s + 1], c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    double ksum = 0.0;
    for (int64_t s = 0; s < std::min(nk, n2); ++s) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]));
        ksum += ms;
    }
    float whole = 0;
    HIP_TRY(hipEventElapsedTime(&whole, ev[2 * nk], ev[2 * nk + 1]));
    if (avg_kernel_ms) *avg_kernel_ms = n2 > 0 ? ksum / double(std::min(nk, n2)) : double(whole) / double(n1);
    if (avg_step_ms) *avg_step_ms = double(whole) / double(n1);
    if (c->d.stats_every > 0) {
        const int64_t k = c->d.stats_every;
        c->stats_steps += (c->steps + n) / k - c->steps / k;
    }
    c->steps += n;
    c->macro_valid_step = c->steps - 1;
    c->strain_valid = false;
    return LBM355_OK;
}

int lbm355_copy_bandwidth(int64_t bytes, int iters, double* gbps)
{
    if (bytes < 1024 || iters < 1 || !gbps) return fail(LBM355_EINVAL, "bad argument");
    bytes &= ~int64_t(15);
    DevBuf a, b;
    if (a.alloc(size_t(bytes)) || b.alloc(size_t(bytes))) return fail(LBM355_ENOMEM, "copy buffers");
    StreamHolder sh;
    HIP_TRY(hipStreamCreateWithFlags(&sh.s, hipStreamNonBlocking));
    hipStream_t st = sh.s;
    EventSet es(2);
    HIP_TRY(hipEventCreate(&es.ev[0]));
    HIP_TRY(hipEventCreate(&es.ev[1]));
    hipEvent_t e0 = es.ev[0], e1 = es.ev[1];
    LAUNCH_TRY(launch_zero(a.p, bytes, st));
    LAUNCH_TRY(launch_copy(b.p, a.p, bytes, st));  // warm-up
    double best = 0.0;
    for (int i = 0; i < iters; ++i) {
        HIP_TRY(hipEventRecord(e0, st));
        LAUNCH_TRY(launch_copy(b.p, a.p, bytes, st));
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (ms > 0) best = std::max(best, 2.0 * double(bytes) / (double(ms) * 1e6));
    }
    *gbps = best;
    return LBM355_OK;
}

int lbm355_case_get_field(lbm355_case* c, int field, double* host, size_t bytes)
{
    if (!c || !host) return fail(LBM355_EINVAL, "null argument");
    const int64_t N = int64_t(c->d.nx) * c->d.ny * c->d.nz;
    if (bytes < size_t(N) * 8) return fail(LBM355_EINVAL, "buffer too small");
    const bool f64 = c->storage == LBM355_F64;
    int rc;
    switch (field) {
    case LBM355_FIELD_RHO: case LBM355_FIELD_VX: case LBM355_FIELD_VY: case LBM355_FIELD_VZ:
        if (c->steps == 0) {  // the initial state: rho = 1, u = 0 everywhere
            std::fill(host, host + N, field == LBM355_FIELD_RHO ? 1.0 : 0.0);
            return LBM355_OK;
        }
        if ((rc = refresh_macro(*c))) return rc;
        return f64 ? gather_dense<double>(*c, c->macro.p, 4, field, field == 0 ? 1.0 : 0.0, host)
                   : gather_dense<float>(*c, c->macro.p, 4, field, field == 0 ? 1.0 : 0.0, host);
    case LBM355_FIELD_FX: case LBM355_FIELD_FY: case LBM355_FIELD_FZ:
        if (!c->force.p) {
            std::fill(host, host + N, 0.0);
            return LBM355_OK;
        }
        return f64 ? gather_dense<double>(*c, c->force.p, 3, field - LBM355_FIELD_FX, 0.0, host)
                   : gather_dense<float>(*c, c->force.p, 3, field - LBM355_FIELD_FX, 0.0, host);
    case LBM355_FIELD_LABEL:
        for (int64_t i = 0; i < N; ++i) host[i] = c->lab.map[i];
        return LBM355_OK;
    case LBM355_FIELD_WALLMASK:
        for (int64_t i = 0; i < N; ++i) host[i] = c->lab.wall[i];
        return LBM355_OK;
    default: break;
    }
    if (field >= LBM355_FIELD_SUM_RUX && field <= LBM355_FIELD_SUM_RRUZZ) {
        if (c->d.stats_every <= 0) { std::fill(host, host + N, 0.0); return LBM355_OK; }
        const int sf = field - LBM355_FIELD_SUM_RUX;
        if ((rc = gather_dense<double>(*c, c->stats.p, kStatFields, sf, 0.0, host))) return rc;
        // samples of the open group are in the ring, not in the FP64 sums yet
        const int64_t open = f64 ? 0 : (c->stats_steps - c->stats_block0) % kStatBlock;
        std::vector<double> part(open > 0 ? static_cast<size_t>(N) : 0);
        for (int64_t sl = 0; sl < open; ++sl) {
            if ((rc = gather_dense<float>(*c, ring_slot(*c, sl), kRingFields, sf % kRingFields, 0.0, part.data())))
                return rc;
            for (int64_t i = 0; i < N; ++i) host[i] += sf < kRingFields ? part[size_t(i)] : part[size_t(i)] * part[size_t(i)];
        }
        return LBM355_OK;
    }
    if (field >= LBM355_FIELD_S11 && field <= LBM355_FIELD_S33) {
        return f64 ? gather_dense<double>(*c, c->strain.p, 6, field - LBM355_FIELD_S11, 0.0, host)
                   : gather_dense<float>(*c, c->strain.p, 6, field - LBM355_FIELD_S11, 0.0, host);
    }
    return fail(LBM355_EINVAL, "unknown field id");
}

int lbm355_case_get_populations(lbm355_case* c, double* host, size_t bytes)
{
    if (!c || !host) return fail(LBM355_EINVAL, "null argument");
    const int64_t N = int64_t(c->d.nx) * c->d.ny * c->d.nz;
    if (bytes < size_t(N) * kQ * 8) return fail(LBM355_EINVAL, "buffer too small");
    const size_t sb = storage_bytes(c->storage);
    const size_t ne = size_t(c->n_stored + 1) * kQ * kLanes;
    std::vector<unsigned char> raw(ne * sb);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(raw.data(), c->f.p, raw.size(), hipMemcpyDeviceToHost));
    auto value = [&](int32_t off, int q) -> double {
        switch (c->storage) {
        case LBM355_F64: { double v; std::memcpy(&v, &raw[size_t(off) * 8], 8); return v; }
        case LBM355_F32: { float v; std::memcpy(&v, &raw[size_t(off) * 4], 4); return double(v) + lattice_weight(q); }
        case LBM355_F16_OFFSET: {
            uint16_t h; std::memcpy(&h, &raw[size_t(off) * 2], 2);
            const uint32_t sgn = uint32_t(h & 0x8000) << 16;
            int e = (h >> 10) & 0x1F;
            uint32_t m = h & 0x3FF, bits;
            if (e == 0) {
                if (m == 0) bits = sgn;
                else { e = 1; while (!(m & 0x400)) { m <<= 1; --e; } m &= 0x3FF; bits = sgn | uint32_t(e + 112) << 23 | m << 13; }
            } else if (e == 31) bits = sgn | 0x7F800000u | m << 13;
            else bits = sgn | uint32_t(e + 112) << 23 | m << 13;
            float v; std::memcpy(&v, &bits, 4);
            return double(std::fma(v, 1.0f / float(1 << half_scale_exp(q)), float(lattice_weight(q))));
        }
        default: {
            uint16_t h; std::memcpy(&h, &raw[size_t(off) * 2], 2);
            const uint32_t bits = uint32_t(h) << 16;
            float v; std::memcpy(&v, &bits, 4);
            return double(v + float(lattice_weight(q)));
        }
        }
    };
    std::fill(host, host + N * kQ, 0.0);
    for_each_held(*c, [&](int32_t off, int q, int64_t k) { host[k] = value(off, q); });
    return LBM355_OK;
}

int lbm355_case_compute_strain(lbm355_case* c)
{
    if (!c) return fail(LBM355_EINVAL, "null case");
    int rc;
    if ((rc = refresh_macro(*c))) return rc;
    LAUNCH_TRY(launch_strain(c->storage, static_cast<const CaseDev*>(c->devcase.p), c->n_active, c->stream));
    c->strain_valid = true;
    return LBM355_OK;
}

int lbm355_case_flux(lbm355_case* c, int mean, double out[4])
{
    if (!c || !out) return fail(LBM355_EINVAL, "null argument");
    if (mean && c->d.stats_every <= 0) return fail(LBM355_EINVAL, "statistics are off");
    // samples of the open group are still in the ring, not in the FP64 sums
    const int open = mean && c->storage != LBM355_F64 ? int((c->stats_steps - c->stats_block0) % kStatBlock) : 0;
    LAUNCH_TRY(launch_zero(c->flux.p, 32, c->stream));
    LAUNCH_TRY(launch_flux(c->storage, static_cast<const CaseDev*>(c->devcase.p), c->n_bc,
                           static_cast<double*>(c->flux.p), mean, open, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(out, c->flux.p, 32, hipMemcpyDeviceToHost));
    if (mean) {
        // density weighted, as the mean_velocity output (sim_tcpc.cu:724-727)
        const double diff = std::max<double>(1.0, double(c->stats_steps));
        for (int p = 0; p < 4; ++p) out[p] /= diff;
    }
    return LBM355_OK;
}

}  // extern "C"

namespace {
// compare needs the same brick index on both sides: the same grid, labels,
// origin shift and active bricks (cases made from one geometry and face table)
int same_layout(const lbm355_case& a, const lbm355_case& b)
{
    const bool same = a.d.nx == b.d.nx && a.d.ny == b.d.ny && a.d.nz == b.d.nz && a.ox == b.ox && a.oy == b.oy &&
                      a.n_active == b.n_active && a.lab.map == b.lab.map &&
                      std::equal(a.brick_xyz.begin(), a.brick_xyz.begin() + 3 * int64_t(a.n_active), b.brick_xyz.begin());
    return same ? LBM355_OK : fail(LBM355_EINVAL, "cases differ in grid, labels or brick index (compare needs one geometry)");
}

// the two passes of the difference norms on a's stream (after ref's queued
// work), partials holding at least 4 * compare_blocks + 4 doubles
int compare_into(lbm355_case* a, lbm355_case* ref, double* partials, double out[4])
{
    int rc;
    if ((rc = same_layout(*a, *ref))) return rc;
    if ((rc = refresh_macro(*a)) || (rc = refresh_macro(*ref))) return rc;
    if (a->steps == 0 || ref->steps == 0) return fail(LBM355_EINVAL, "compare needs a step on both cases");
    HIP_TRY(hipEventRecord(ref->joined, ref->stream));
    HIP_TRY(hipStreamWaitEvent(a->stream, ref->joined, 0));
    LAUNCH_TRY(launch_compare(a->storage, ref->storage, static_cast<const CaseDev*>(a->devcase.p),
                              static_cast<const CaseDev*>(ref->devcase.p), a->n_active, partials, a->stream));
    const int nb = compare_blocks(a->n_active);
    HIP_TRY(hipMemcpyAsync(out, partials + 4 * nb, 4 * sizeof(double), hipMemcpyDeviceToHost, a->stream));
    HIP_TRY(hipStreamSynchronize(a->stream));
    return LBM355_OK;
}
}  // namespace

extern "C" {

int lbm355_case_compare(lbm355_case* a, lbm355_case* ref, double out[4])
{
    if (!a || !ref || !out) return fail(LBM355_EINVAL, "null argument");
    DevBuf partials;
    if (partials.alloc(size_t(4 * compare_blocks(a->n_active) + 4) * sizeof(double)) != 0)
        return fail(LBM355_ENOMEM, "compare partials");
    return compare_into(a, ref, static_cast<double*>(partials.p), out);
}

int lbm355_case_compare_noalloc(lbm355_case* a, lbm355_case* ref, double out[4])
{
    if (!a || !ref || !out) return fail(LBM355_EINVAL, "null argument");
    if (a->device != ref->device) return fail(LBM355_EINVAL, "cases on different devices");
    const int nb = compare_blocks(a->n_active);
    double* ptr = nullptr;
    int max_blocks = 0;
    if (compare_scratch(a->device, &ptr, &max_blocks)) return fail(LBM355_EDEVICE, kernel_error());
    if (nb > max_blocks)
        return fail(LBM355_EINVAL, "case has " + std::to_string(a->n_active) + " active bricks; the static scratch of "
                                   "the allocation-free compare holds " + std::to_string(4 * max_blocks) +
                                   " (use lbm355_case_compare)");
    return compare_into(a, ref, ptr, out);
}

int lbm355_case_reset_stats(lbm355_case* c)
{
    if (!c) return fail(LBM355_EINVAL, "null case");
    if (c->d.stats_every > 0) {
        LAUNCH_TRY(launch_zero(c->stats.p, int64_t(c->stats.bytes), c->stream));
        int rc;
        if ((rc = set_case_u64(*c, &c->host.stats_epoch, uint64_t(c->steps)))) return rc;
    }
    c->stats_steps = 0;
    c->stats_block0 = 0;
    c->stats_first = c->steps;
    return LBM355_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// checkpoint / restart in the reference's current_state layout
namespace {

const int kCounters = 11;  // STAT_RESET ... VTK3DCUT (state.h:72)
const int kMacroTcpc = 20; // MacroTcpc::N (sim_tcpc.cu:328-335)

std::string state_file(const char* dir, const char* name)
{
    return std::string(dir) + "/" + name + "_rank000";
}

int make_dirs(const std::string& path)
{
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty()) ::mkdir(cur.c_str(), 0777);
        }
        if (i < path.size()) cur.push_back(path[i]);
    }
    struct stat st;
    return ::stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode) ? 0 : -1;
}

template <class T>
int write_raw(const std::string& f, const std::vector<T>& v)
{
    FILE* fp = std::fopen(f.c_str(), "wb");
    if (!fp) return fail(LBM355_EINVAL, "cannot create " + f);
    const size_t n = std::fwrite(v.data(), sizeof(T), v.size(), fp);
    std::fclose(fp);
    return n == v.size() ? LBM355_OK : fail(LBM355_EINVAL, "short write " + f);
}

template <class T>
int read_raw(const std::string& f, std::vector<T>& v)
{
    FILE* fp = std::fopen(f.c_str(), "rb");
    if (!fp) return fail(LBM355_EINVAL, "cannot open " + f);
    const size_t n = std::fread(v.data(), sizeof(T), v.size(), fp);
    std::fclose(fp);
    return n == v.size() ? LBM355_OK : fail(LBM355_EINVAL, "short read " + f);
}

// macro field ids of MacroTcpc -> our dense fields (-1: zero; e_vm_minus_x is
// kept inside e_vm_plus_x, only their sum is read, sim_tcpc.cu:724-725)
const int kMacroMap[kMacroTcpc] = {
    LBM355_FIELD_RHO, LBM355_FIELD_VX, LBM355_FIELD_VY, LBM355_FIELD_VZ,
    LBM355_FIELD_SUM_RUX, -1, LBM355_FIELD_SUM_RUY, LBM355_FIELD_SUM_RUZ,
    LBM355_FIELD_SUM_RRUXX, LBM355_FIELD_SUM_RRUYY, LBM355_FIELD_SUM_RRUZZ,
    LBM355_FIELD_FX, LBM355_FIELD_FY, LBM355_FIELD_FZ,
    LBM355_FIELD_S11, LBM355_FIELD_S12, LBM355_FIELD_S13, LBM355_FIELD_S22, LBM355_FIELD_S32, LBM355_FIELD_S33};

template <class T>
int save_arrays(lbm355_case* c, const char* dir, const std::vector<double>& pops, const std::vector<float>& dev)
{
    const int64_t X = c->d.nx, Y = c->d.ny, Z = c->d.nz, N = X * Y * Z;
    // df: the reference's lat layout x, q, z, y (y fastest, defs.h:57-58)
    std::vector<T> df(size_t(N) * kQ);
    size_t k = 0;
    for (int64_t x = 0; x < X; ++x)
        for (int q = 0; q < kQ; ++q)
            for (int64_t z = 0; z < Z; ++z)
                for (int64_t y = 0; y < Y; ++y) df[k++] = T(pops[size_t(q * N + x + X * (y + Y * z))]);
    int rc;
    // both buffers hold the current populations; the one the next step
    // reads is df_(iterations % 2)
    if ((rc = write_raw(state_file(dir, "df_0"), df))) return rc;
    if ((rc = write_raw(state_file(dir, "df_1"), df))) return rc;
    if (!dev.empty()) {
        // an FP32 case's stored words f - w (f32_words), in the same layout,
        // for a bit-exact restart; a reader of the reference layout ignores
        // the extra file
        std::vector<float> out(size_t(N) * kQ);
        k = 0;
        for (int64_t x = 0; x < X; ++x)
            for (int q = 0; q < kQ; ++q)
                for (int64_t z = 0; z < Z; ++z)
                    for (int64_t y = 0; y < Y; ++y) out[k++] = dev[size_t(q * N + x + X * (y + Y * z))];
        if ((rc = write_raw(state_file(dir, "df_dev"), out))) return rc;
    }
    std::vector<int16_t> map(static_cast<size_t>(N));
    k = 0;
    for (int64_t x = 0; x < X; ++x)
        for (int64_t z = 0; z < Z; ++z)
            for (int64_t y = 0; y < Y; ++y) map[k++] = int16_t(c->lab.map[size_t(x + X * (y + Y * z))]);
    if ((rc = write_raw(state_file(dir, "map"), map))) return rc;
    std::vector<double> f(static_cast<size_t>(N));
    std::vector<T> macro(static_cast<size_t>(N) * kMacroTcpc, T(0));
    for (int id = 0; id < kMacroTcpc; ++id) {
        if (kMacroMap[id] < 0) continue;
        if ((rc = lbm355_case_get_field(c, kMacroMap[id], f.data(), f.size() * 8))) return rc;
        for (int64_t x = 0; x < X; ++x)
            for (int64_t z = 0; z < Z; ++z)
                for (int64_t y = 0; y < Y; ++y)
                    macro[size_t(y + Y * (z + Z * (id + int64_t(kMacroTcpc) * x)))] = T(f[size_t(x + X * (y + Y * z))]);
    }
    return write_raw(state_file(dir, "macro"), macro);
}

template <class T>
int load_arrays(lbm355_case* c, const char* dir, int64_t iterations, std::vector<double>& pops,
                std::vector<std::vector<double>>& fields, std::vector<float>& dev)
{
    const int64_t X = c->d.nx, Y = c->d.ny, Z = c->d.nz, N = X * Y * Z;
    std::vector<T> df(size_t(N) * kQ);
    int rc;
    if ((rc = read_raw(state_file(dir, iterations % 2 ? "df_1" : "df_0"), df))) return rc;
    pops.assign(size_t(N) * kQ, 0.0);
    size_t k = 0;
    for (int64_t x = 0; x < X; ++x)
        for (int q = 0; q < kQ; ++q)
            for (int64_t z = 0; z < Z; ++z)
                for (int64_t y = 0; y < Y; ++y) pops[size_t(q * N + x + X * (y + Y * z))] = double(df[k++]);
    // an FP32 case's own checkpoint also holds its stored words f - w
    // (save_arrays); a reference checkpoint has only f, from which the
    // deviations follow to FP32 rounding of f
    dev.clear();
    const std::string devf = state_file(dir, "df_dev");
    struct stat dst;
    if (c->storage == LBM355_F32 && ::stat(devf.c_str(), &dst) == 0) {
        std::vector<float> in(size_t(N) * kQ);
        if ((rc = read_raw(devf, in))) return rc;
        dev.assign(size_t(N) * kQ, 0.0f);
        k = 0;
        for (int64_t x = 0; x < X; ++x)
            for (int q = 0; q < kQ; ++q)
                for (int64_t z = 0; z < Z; ++z)
                    for (int64_t y = 0; y < Y; ++y) dev[size_t(q * N + x + X * (y + Y * z))] = in[k++];
    }
    std::vector<T> macro(size_t(N) * kMacroTcpc);
    if ((rc = read_raw(state_file(dir, "macro"), macro))) return rc;
    fields.assign(kMacroTcpc, std::vector<double>());
    for (int id = 0; id < kMacroTcpc; ++id) {
        fields[id].assign(size_t(N), 0.0);
        for (int64_t x = 0; x < X; ++x)
            for (int64_t z = 0; z < Z; ++z)
                for (int64_t y = 0; y < Y; ++y)
                    fields[id][size_t(x + X * (y + Y * z))] =
                        double(macro[size_t(y + Y * (z + Z * (id + int64_t(kMacroTcpc) * x)))]);
    }
    // the AA-split x statistic is their sum
    for (int64_t i = 0; i < N; ++i) fields[4][size_t(i)] += fields[5][size_t(i)];
    return LBM355_OK;
}

uint16_t host_f16(float v)
{
    const _Float16 h = static_cast<_Float16>(v);  // round to nearest even
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}

uint16_t host_bf16(float v)
{
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

// write dense A/B-sense populations (as lbm355_case_get_populations returns
// them) into the AA buffer for the case's current step parity
int set_populations(lbm355_case* c, const std::vector<double>& pops)
{
    const size_t sb = storage_bytes(c->storage);
    const size_t ne = size_t(c->n_stored + 1) * kQ * kLanes;
    std::vector<unsigned char> raw(ne * sb);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(raw.data(), c->f.p, raw.size(), hipMemcpyDeviceToHost));
    auto put = [&](int32_t off, int q, double v) {
        switch (c->storage) {
        case LBM355_F64: std::memcpy(&raw[size_t(off) * 8], &v, 8); break;
        case LBM355_F32: { const float f = float(v - lattice_weight(q)); std::memcpy(&raw[size_t(off) * 4], &f, 4); break; }
        case LBM355_F16_OFFSET: {
            const uint16_t h = host_f16((float(v) - float(lattice_weight(q))) * float(1 << half_scale_exp(q)));
            std::memcpy(&raw[size_t(off) * 2], &h, 2);
            break;
        }
        default: {
            const uint16_t h = host_bf16(float(v) - float(lattice_weight(q)));
            std::memcpy(&raw[size_t(off) * 2], &h, 2);
        }
        }
    };
    for_each_held(*c, [&](int32_t off, int q, int64_t k) { put(off, q, pops[size_t(k)]); });
    HIP_TRY(hipMemcpy(c->f.p, raw.data(), raw.size(), hipMemcpyHostToDevice));
    return LBM355_OK;
}

// the stored FP32 words (f - w) of the held slots, dense [q][cell], read
// (Write = false) or written back.  An FP32 checkpoint keeps them beside the
// reference layout's float f: w + d is not exact in a double once d is small
// (d = 1e-11 has bits below the double spacing near w), so the dense doubles
// of lbm355_case_get_populations cannot carry them either
template <bool Write>
int f32_words(lbm355_case* c, std::vector<float>& dev)
{
    const int64_t N = int64_t(c->d.nx) * c->d.ny * c->d.nz;
    std::vector<float> raw(size_t(c->n_stored + 1) * kQ * kLanes);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(raw.data(), c->f.p, raw.size() * 4, hipMemcpyDeviceToHost));
    if (!Write) dev.assign(size_t(N) * kQ, 0.0f);
    for_each_held(*c, [&](int32_t off, int, int64_t k) {
        if (Write) raw[size_t(off)] = dev[size_t(k)];
        else dev[size_t(k)] = raw[size_t(off)];
    });
    if (Write) HIP_TRY(hipMemcpy(c->f.p, raw.data(), raw.size() * 4, hipMemcpyHostToDevice));
    return LBM355_OK;
}

}  // namespace

extern "C" {

int lbm355_case_save_state(lbm355_case* c, const char* dir, const lbm355_state_desc* s)
{
    if (!c || !dir || !s) return fail(LBM355_EINVAL, "null argument");
    if (make_dirs(dir) != 0) return fail(LBM355_EINVAL, std::string("cannot create ") + dir);
    const int64_t X = c->d.nx, Y = c->d.ny, Z = c->d.nz, N = X * Y * Z;
    std::vector<double> pops(size_t(N) * kQ);
    int rc;
    if ((rc = lbm355_case_get_populations(c, pops.data(), pops.size() * 8))) return rc;
    // config and counters as saveLoadTextData writes them (state.hpp:1478-1500, state.h:236-250)
    FILE* fp = std::fopen(state_file(dir, "config").c_str(), "wt");
    if (!fp) return fail(LBM355_EINVAL, "cannot create config");
    std::fprintf(fp, "%d\n%ld\n%ld\n%ld\n%ld\n%ld\n%ld\n%le\n", int(c->steps), long(X), long(Y), long(Z), long(X),
                 long(Y), long(Z), s->phys_final_time);
    std::fclose(fp);
    for (int k = 0; k < kCounters; ++k) {
        const std::string name = "cnt_" + std::to_string(k);
        fp = std::fopen(state_file(dir, name.c_str()).c_str(), "wt");
        if (!fp) return fail(LBM355_EINVAL, "cannot create " + name);
        std::fprintf(fp, "%d\n%le\n", s->counter_count[k], s->counter_period[k]);
        std::fclose(fp);
    }
    // not in the reference layout: what an exact restart of this library needs
    fp = std::fopen(state_file(dir, "lbm355").c_str(), "wt");
    if (!fp) return fail(LBM355_EINVAL, "cannot create lbm355 state");
    std::fprintf(fp, "%d\n%d\n%ld\n%ld\n", c->storage, c->model, long(s->stat_prev_iter), long(c->stats_steps));
    std::fclose(fp);
    std::vector<float> dev;
    if (c->storage == LBM355_F32 && (rc = f32_words<false>(c, dev))) return rc;
    return c->storage == LBM355_F64 ? save_arrays<double>(c, dir, pops, dev) : save_arrays<float>(c, dir, pops, dev);
}

int lbm355_case_load_state(lbm355_case* c, const char* dir, lbm355_state_desc* s)
{
    if (!c || !dir || !s) return fail(LBM355_EINVAL, "null argument");
    FILE* fp = std::fopen(state_file(dir, "config").c_str(), "rt");
    if (!fp) return fail(LBM355_EINVAL, std::string("no saved state in ") + dir);
    int it = 0;
    long gx = 0, gy = 0, gz = 0, lx = 0, ly = 0, lz = 0;
    double ft = 0.0;
    const int nr = std::fscanf(fp, "%d %ld %ld %ld %ld %ld %ld %le", &it, &gx, &gy, &gz, &lx, &ly, &lz, &ft);
    std::fclose(fp);
    if (nr != 8) return fail(LBM355_EINVAL, "malformed config");
    if (gx != c->d.nx || gy != c->d.ny || gz != c->d.nz) return fail(LBM355_EINVAL, "saved state has other dimensions");
    s->phys_final_time = ft;
    s->iterations = it;
    for (int k = 0; k < kCounters; ++k) {
        const std::string name = "cnt_" + std::to_string(k);
        fp = std::fopen(state_file(dir, name.c_str()).c_str(), "rt");
        if (!fp) return fail(LBM355_EINVAL, "missing " + name);
        int cnt = 0;
        double per = -1.0;
        const int r = std::fscanf(fp, "%d %le", &cnt, &per);
        std::fclose(fp);
        if (r != 2) return fail(LBM355_EINVAL, "malformed " + name);
        s->counter_count[k] = cnt;
        s->counter_period[k] = per;
    }
    long prev = 0, samples = -1;
    int st = -1, mo = -1;
    if ((fp = std::fopen(state_file(dir, "lbm355").c_str(), "rt"))) {
        if (std::fscanf(fp, "%d %d %ld %ld", &st, &mo, &prev, &samples) != 4) samples = -1;
        std::fclose(fp);
    }
    s->stat_prev_iter = prev;
    std::vector<double> pops;
    std::vector<std::vector<double>> fields;
    int rc;
    // a reference state is dreal = float unless this case stores FP64
    const bool f64 = c->storage == LBM355_F64;
    std::vector<float> dev;
    rc = f64 ? load_arrays<double>(c, dir, it, pops, fields, dev) : load_arrays<float>(c, dir, it, pops, fields, dev);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->steps = it;
    if ((rc = set_populations(c, pops))) return rc;
    if (!dev.empty() && (rc = f32_words<true>(c, dev))) return rc;
    // macro rho,u of the saved step, statistics sums, body force, strain
    for (int id = 0; id < kMacroTcpc; ++id) {
        const int f = kMacroMap[id];
        if (f < 0) continue;
        if (f <= LBM355_FIELD_VZ)
            rc = f64 ? scatter_dense<double>(*c, c->macro.p, 4, f, fields[id].data())
                     : scatter_dense<float>(*c, c->macro.p, 4, f, fields[id].data());
        else if (f >= LBM355_FIELD_SUM_RUX && f <= LBM355_FIELD_SUM_RRUZZ) {
            if (c->d.stats_every > 0) rc = scatter_dense<double>(*c, c->stats.p, kStatFields, f - LBM355_FIELD_SUM_RUX,
                                                                fields[id].data());
        } else if (f >= LBM355_FIELD_FX && f <= LBM355_FIELD_FZ) {
            if (c->force.p)
                rc = f64 ? scatter_dense<double>(*c, c->force.p, 3, f - LBM355_FIELD_FX, fields[id].data())
                         : scatter_dense<float>(*c, c->force.p, 3, f - LBM355_FIELD_FX, fields[id].data());
        } else if (f >= LBM355_FIELD_S11 && f <= LBM355_FIELD_S33) {
            rc = f64 ? scatter_dense<double>(*c, c->strain.p, 6, f - LBM355_FIELD_S11, fields[id].data())
                     : scatter_dense<float>(*c, c->strain.p, 6, f - LBM355_FIELD_S11, fields[id].data());
        }
        if (rc) return rc;
    }
    // statistics: samples since the last reset; sample groups restart here
    const int64_t k = c->d.stats_every > 0 ? c->d.stats_every : 1;
    c->stats_steps = samples >= 0 ? samples : int64_t(it) / k - int64_t(prev) / k;
    c->stats_block0 = c->stats_steps;
    c->stats_first = c->steps;
    if (c->d.stats_every > 0 && (rc = set_case_u64(*c, &c->host.stats_epoch, uint64_t(c->steps)))) return rc;
    c->macro_valid_step = c->steps - 1;
    c->strain_valid = false;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM355_OK;
}

int lbm355_case_write_vtk(lbm355_case* c, const char* path, const lbm355_output_desc* o)
{
    if (!c || !path || !o) return fail(LBM355_EINVAL, "null argument");
    const int X = c->d.nx, Y = c->d.ny, Z = c->d.nz;
    const int64_t N = int64_t(X) * Y * Z;
    const double dl = c->d.dl, dt = c->d.dt;
    int rc;
    if ((rc = lbm355_case_compute_strain(c))) return rc;
    std::vector<double> F[16];
    const int ids[16] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};
    for (int k = 0; k < 16; ++k) {
        F[k].resize(N);
        if ((rc = lbm355_case_get_field(c, ids[k], F[k].data(), size_t(N) * 8))) return rc;
    }
    const std::vector<double>& rho = F[0];
    const double vconv = dl / dt;
    const double diff = std::max(1.0, double(o->stat_steps));
    const double cs2 = vconv * vconv / 3.0;
    const bool cy = c->d.cy_enabled != 0 || o->material_cy != 0;
    auto wall_at = [&](int x, int y, int z) -> bool {
        if (x < 0 || y < 0 || z < 0 || x >= X || y >= Y || z >= Z) return false;
        return c->lab.wall[c->cell_index(x, y, z)] != 0;
    };
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(LBM355_EINVAL, std::string("cannot open ") + path);
    std::vector<char> buf;
    auto flush = [&]() { if (!buf.empty()) std::fwrite(buf.data(), 1, buf.size(), fp); buf.clear(); };
    std::fprintf(fp, "# vtk DataFile Version 3.0\nWritten using RF writer\nBINARY\n");
    std::fprintf(fp, "DATASET RECTILINEAR_GRID\nDIMENSIONS %d %d %d\n", X, Y, Z);
    std::fprintf(fp, "X_COORDINATES %d float\n", X);
    for (int x = 0; x < X; ++x) put_float(buf, float(x * dl));
    flush();
    std::fprintf(fp, "Y_COORDINATES %d float\n", Y);
    for (int y = 0; y < Y; ++y) put_float(buf, float(y * dl));
    flush();
    std::fprintf(fp, "Z_COORDINATES %d float\n", Z);
    for (int z = 0; z < Z; ++z) put_float(buf, float(z * dl));
    flush();
    std::fprintf(fp, "FIELD FieldData %d\n", 2);
    std::fprintf(fp, "TIME %d %d float\n", 1, 1);
    put_float(buf, float(o->time));
    flush();
    std::fprintf(fp, "CYCLE %d %d float\n", 1, 1);
    put_float(buf, float(o->cycle));
    flush();
    std::fprintf(fp, "POINT_DATA %d\n", int(N));
    std::fprintf(fp, "SCALARS wall int 1\nLOOKUP_TABLE default\n");
    for (int64_t i = 0; i < N; ++i) put_be32(buf, uint32_t(int32_t(c->lab.map[i])));
    flush();
    // StateTcpc::outputData, sim_tcpc.cu:716-784
    auto local_nu = [&](double g) {
        if (!cy) return c->d.nu_ref;
        return c->d.nu_ref + (c->d.cy_nu0 - c->d.nu_ref) * std::pow(1.0 + std::pow(g * c->d.cy_lambda, c->d.cy_a),
                                                                   (c->d.cy_n - 1.0) / c->d.cy_a);
    };
    auto gamma_lbm = [&](int64_t i) {
        const double s11 = F[10][i], s12 = F[11][i], s13 = F[12][i], s22 = F[13][i], s32 = F[14][i], s33 = F[15][i];
        return std::sqrt(2.0) * std::sqrt(s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s32 * s32));
    };
    const char* names[12] = {"velocity", "mean_velocity", "lbm_rho", "pressure_pa", "gamma_dot", "nu_phys_m2_s",
                             "mu_phys_pa_s", "wss_proxy_pa", "turbulence_intensity", "rms_velocity",
                             "strain_diag_1_s", "strain_shear_1_s"};
    const int dofs[12] = {3, 3, 1, 1, 1, 1, 1, 1, 1, 1, 3, 3};
    for (int a = 0; a < 12; ++a) {
        if (dofs[a] == 1) std::fprintf(fp, "SCALARS %s float 1\nLOOKUP_TABLE default\n", names[a]);
        else std::fprintf(fp, "VECTORS %s float\n", names[a]);
        for (int z = 0; z < Z; ++z)
            for (int y = 0; y < Y; ++y)
                for (int x = 0; x < X; ++x) {
                    const int64_t i = c->cell_index(x, y, z);
                    const double mvx = F[4][i] * vconv / diff, mvy = F[5][i] * vconv / diff, mvz = F[6][i] * vconv / diff;
                    double vals[3] = {0, 0, 0};
                    switch (a) {
                    case 0: vals[0] = F[1][i] * vconv; vals[1] = F[2][i] * vconv; vals[2] = F[3][i] * vconv; break;
                    case 1: vals[0] = mvx; vals[1] = mvy; vals[2] = mvz; break;
                    case 2: vals[0] = rho[i]; break;
                    case 3: vals[0] = (rho[i] - 1.0) * o->phys_density * cs2; break;
                    case 4: vals[0] = gamma_lbm(i) / dt; break;
                    case 5: vals[0] = local_nu(gamma_lbm(i)) * dl * dl / dt; break;
                    case 6: vals[0] = local_nu(gamma_lbm(i)) * dl * dl / dt * o->phys_density; break;
                    case 7: {
                        const bool near = wall_at(x - 1, y, z) || wall_at(x + 1, y, z) || wall_at(x, y - 1, z) ||
                                          wall_at(x, y + 1, z) || wall_at(x, y, z - 1) || wall_at(x, y, z + 1);
                        const double g = gamma_lbm(i);
                        vals[0] = near ? local_nu(g) * dl * dl / dt * o->phys_density * (g / dt) : 0.0;
                        break;
                    }
                    case 8: case 9: {
                        const double mxx = F[7][i] * vconv * vconv / diff, myy = F[8][i] * vconv * vconv / diff,
                                     mzz = F[9][i] * vconv * vconv / diff;
                        const double rx = std::fabs(mxx - mvx * mvx), ry = std::fabs(myy - mvy * mvy),
                                     rz = std::fabs(mzz - mvz * mvz);
                        const double rms = std::sqrt(rx * rx + ry * ry + rz * rz);
                        const double mv = std::sqrt(mvx * mvx + mvy * mvy + mvz * mvz);
                        vals[0] = a == 9 ? rms : (std::fabs(mv) <= 1e-10 ? 0.0 : rms / std::fabs(mv));
                        break;
                    }
                    case 10: vals[0] = F[10][i] / dt; vals[1] = F[13][i] / dt; vals[2] = F[15][i] / dt; break;
                    case 11: vals[0] = F[11][i] / dt; vals[1] = F[12][i] / dt; vals[2] = F[14][i] / dt; break;
                    }
                    for (int k = 0; k < dofs[a]; ++k) put_float(buf, float(vals[k]));
                    if (buf.size() > (1u << 20)) flush();
                }
        flush();
    }
    std::fclose(fp);
    return LBM355_OK;
}

// ---------------------------------------------------------------------------
// batches: several cases of one collision and storage type stepped by one
// launch per step over their concatenated brick lists.  The list, the
// members' constants and the brick range of the group being stepped live in
// device memory (BatchDev), read by the batch kernels at run time: a
// membership change re-uploads them and leaves a captured graph valid unless
// the launch grid has to grow.
struct lbm355_batch {
    std::vector<lbm355_case*> cases;
    int storage = 1, model = 0;
    int use_graph = 0;
    int bricks_per_group = 4;
    int nontemporal = -1;        // -1 auto (by the members' working set), 0 off, 1 on
    int waves_cap = 0;           // wavefronts per CU of the fused step, 0 = no cap
    int per_wave = 1;            // bricks per wavefront: 1 one each (k_step_range), k > 1 the persistent loop
    hipStream_t stream = nullptr;
    // recorded on the batch's stream after its steps; every member's stream
    // waits on it, so whatever is queued for a case later runs after them
    hipEvent_t done = nullptr;
    DevBuf devcases, list, base, desc;
    int64_t cap_cases = 0, cap_bricks = 0;  // capacity of devcases / list (elements)
    int n_list = 0;
    // the cases are stepped in groups of consecutive members whose device
    // memory a step sweeps stays below group_bytes (0: one group); each group
    // runs a whole chunk of steps before the next, so a group that fits the
    // Infinity Cache stays there for the chunk
    struct Group {
        int begin, end;   // brick list range
        int64_t bytes;
    };
    std::vector<Group> groups;
    std::vector<int> case_first;        // brick list offset of each member
    std::vector<int64_t> case_bytes;    // device memory a step of each member sweeps (case_step_bytes)
    int64_t group_bytes = kBatchGroupBytes;
    int64_t launches = 0;     // steps done by the batch; the device launch counter equals it between calls
    int cur_begin = -1, cur_end = -1;   // range last written to the device descriptor
    std::vector<int64_t> start;  // per case: launch index of its step 0 in the batch copy
    std::vector<int64_t> epoch;  // per case: statistics epoch in the batch copy
    std::vector<double> in_xm, in_xp;  // per case: inflow scales in the batch copy
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_grid = 0;          // launch grid baked into the graph (workgroups)
    int64_t uploads = 0, captures = 0;  // membership uploads and graph captures so far
    ~lbm355_batch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) { (void)hipStreamSynchronize(stream); (void)hipStreamDestroy(stream); }
        if (done) (void)hipEventDestroy(done);
    }
};

namespace {
// the autotuner's search space: bricks per workgroup x even-step cache policy
// x wavefronts per CU (no cap, 16, 12)
struct TuneConfig {
    int bpg, nt, cap;
};
constexpr int kTuneConfigs = 4 * 2 * 3;
constexpr int kTuneFinalists = 4;   // configurations timed again on longer runs (single cases)
constexpr int kTunePasses = 3;      // interleaved passes over the finalists, each keeping its fastest
constexpr int kTuneWarmup = 64;     // steps run before any timing (clocks settle)
// timed window of the autotuners: a whole number of statistics groups, so
// every window holds the same number of group flushes (a flush step costs
// about twice a plain one; 8-step windows of a 16-sample group held one flush
// or none); 8 steps without statistics, at most 128
int tune_window(int stats_every)
{
    if (stats_every <= 0) return 8;
    return int(std::min<int64_t>(int64_t(kStatBlock) * stats_every, 128));
}
TuneConfig tune_config(int k)
{
    const int shapes[4] = {1, 2, 4, 8}, caps[3] = {0, 16, 12};
    return TuneConfig{shapes[k % 4], (k / 4) % 2, caps[k / 8]};
}

// split the members into groups of consecutive cases below group_bytes; a
// case larger than that is a group of its own
void regroup(lbm355_batch& b)
{
    b.groups.clear();
    const int n = int(b.cases.size());
    for (int i = 0; i < n; ++i) {
        const int end = i + 1 < n ? b.case_first[i + 1] : b.n_list;
        if (!b.groups.empty() && (b.group_bytes <= 0 || b.groups.back().bytes + b.case_bytes[i] <= b.group_bytes)) {
            b.groups.back().end = end;
            b.groups.back().bytes += b.case_bytes[i];
        } else {
            b.groups.push_back({b.case_first[i], end, b.case_bytes[i]});
        }
    }
}

void drop_graph(lbm355_batch& b)
{
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
    b.graph_grid = 0;
}

// launch grid (workgroups) that covers the largest group: one workgroup of
// bricks_per_group bricks each, or per_wave bricks per wavefront for the
// persistent loop; a multiple of 8 for the XCD split
int batch_grid(const lbm355_batch& b)
{
    int most = 0;
    for (const auto& g : b.groups) most = std::max(most, g.end - g.begin);
    int blocks = (most + b.bricks_per_group - 1) / b.bricks_per_group;
    if (b.per_wave > 1) blocks = (blocks + b.per_wave - 1) / b.per_wave;
    return std::max(8, (blocks + 7) / 8 * 8);
}

StepLaunch batch_step_launch(lbm355_batch& b, int grid)
{
    StepLaunch L{};
    L.batch = static_cast<const BatchDev*>(b.desc.p);
    L.n_list = grid;
    L.base = static_cast<const int64_t*>(b.base.p);
    L.rel = 0;
    L.bricks_per_group = b.bricks_per_group;
    int64_t most = 0;
    for (const auto& g : b.groups) most = std::max(most, g.bytes);
    L.nontemporal = b.nontemporal >= 0 ? b.nontemporal : (most > kNontemporalBytes ? 1 : 0);
    L.lds_bytes = occupancy_lds(b.waves_cap, b.bricks_per_group);
    L.per_wave = b.per_wave;
    return L;
}

// point the device descriptor at brick range [begin, end) (a kernel on the
// batch's stream: it orders with the steps around it and records no memset)
int set_range(lbm355_batch& b, int begin, int end)
{
    if (begin == b.cur_begin && end == b.cur_end) return LBM355_OK;
    BatchDev v;
    v.list = static_cast<const BrickRef*>(b.list.p);
    v.cases = static_cast<CaseDev*>(b.devcases.p);
    v.begin = begin;
    v.end = end;
    LAUNCH_TRY(launch_set_batch(static_cast<BatchDev*>(b.desc.p), v, b.stream));
    b.cur_begin = begin;
    b.cur_end = end;
    return LBM355_OK;
}

// grow a device array to hold `need` elements of `elem` bytes (contents are
// re-uploaded by the caller)
int reserve(DevBuf& buf, int64_t& cap, int64_t need, size_t elem)
{
    if (need <= cap && buf.p) return LBM355_OK;
    const int64_t n = std::max<int64_t>({need + need / 2, 16, cap});
    DevBuf fresh;
    if (fresh.alloc(size_t(n) * elem) != 0) return fail(LBM355_ENOMEM, "batch arrays");
    std::swap(buf.p, fresh.p);
    std::swap(buf.bytes, fresh.bytes);
    cap = n;
    return LBM355_OK;
}

// rebuild the brick list, the member constants and the groups after a
// membership change and upload them; the batch's stream is idle.  A member's
// step 0 sits at launch index launches - steps, so its own step count goes on.
int batch_upload(lbm355_batch& b)
{
    std::vector<CaseDev> cd;
    std::vector<BrickRef> list;
    b.start.clear(); b.epoch.clear(); b.in_xm.clear(); b.in_xp.clear();
    b.case_first.clear(); b.case_bytes.clear();
    for (size_t i = 0; i < b.cases.size(); ++i) {
        lbm355_case* c = b.cases[i];
        cd.push_back(c->host);
        CaseDev& h = cd.back();
        h.start = b.launches - c->steps;
        h.stats_epoch = c->stats_first;
        h.macro_step = -1;  // macro output only on explicit request inside a batch
        b.start.push_back(h.start);
        b.epoch.push_back(h.stats_epoch);
        b.in_xm.push_back(h.inflow_xm);
        b.in_xp.push_back(h.inflow_xp);
        b.case_first.push_back(int(list.size()));
        b.case_bytes.push_back(case_step_bytes(*c));
        for (int k = 0; k < c->n_active; ++k) list.push_back(BrickRef{uint32_t(i), uint32_t(k)});
    }
    // the cases' slab-ordered brick runs follow each other; the
    // XCD-contiguous split of a group's range gives each XCD consecutive bricks
    b.n_list = int(list.size());
    regroup(b);
    int rc;
    if ((rc = reserve(b.devcases, b.cap_cases, int64_t(cd.size()), sizeof(CaseDev)))) return rc;
    if ((rc = reserve(b.list, b.cap_bricks, int64_t(list.size()), sizeof(BrickRef)))) return rc;
    if (!cd.empty()) HIP_TRY(hipMemcpy(b.devcases.p, cd.data(), cd.size() * sizeof(CaseDev), hipMemcpyHostToDevice));
    if (!list.empty())
        HIP_TRY(hipMemcpy(b.list.p, list.data(), list.size() * sizeof(BrickRef), hipMemcpyHostToDevice));
    b.cur_begin = b.cur_end = -1;  // the descriptor's pointers may have moved
    if ((rc = set_range(b, 0, 0))) return rc;
    HIP_TRY(hipStreamSynchronize(b.stream));
    // the graph bakes the grid; a larger group needs a larger one
    if (b.exec && batch_grid(b) > b.graph_grid) drop_graph(b);
    ++b.uploads;
    return LBM355_OK;
}

int check_member(const lbm355_batch& b, const lbm355_case* c)
{
    if (!c) return fail(LBM355_EINVAL, "null case in batch");
    if (c->storage != b.storage || c->model != b.model)
        return fail(LBM355_EINVAL, "a batch needs one collision and one storage type");
    if (c->host.force) return fail(LBM355_EINVAL, "fd-force cases step alone (their pre-kernels are per case)");
    if (std::find(b.cases.begin(), b.cases.end(), c) != b.cases.end())
        return fail(LBM355_EINVAL, "case is already a member of the batch");
    return LBM355_OK;
}
}  // namespace

int lbm355_batch_create(lbm355_case* const* cases, int n, int graph, lbm355_batch** out)
{
    if (!cases || n <= 0 || !out) return fail(LBM355_EINVAL, "bad batch arguments");
    std::unique_ptr<lbm355_batch> b(new (std::nothrow) lbm355_batch());
    if (!b) return fail(LBM355_ENOMEM, "host allocation");
    if (!cases[0]) return fail(LBM355_EINVAL, "null case in batch");
    b->storage = cases[0]->storage;
    b->model = cases[0]->model;
    b->use_graph = graph;
    int rc;
    for (int i = 0; i < n; ++i) {
        if ((rc = check_member(*b, cases[i]))) return rc;
        HIP_TRY(hipStreamSynchronize(cases[i]->stream));
        b->cases.push_back(cases[i]);
    }
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    if (b->base.alloc(16) != 0 || b->desc.alloc(sizeof(BatchDev)) != 0) return fail(LBM355_ENOMEM, "batch counters");
    LAUNCH_TRY(launch_zero(b->base.p, 16, b->stream));
    if ((rc = batch_upload(*b))) return rc;
    *out = b.release();
    return LBM355_OK;
}

int lbm355_batch_add(lbm355_batch* b, lbm355_case* c)
{
    if (!b) return fail(LBM355_EINVAL, "null batch");
    int rc;
    if ((rc = check_member(*b, c))) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    b->cases.push_back(c);
    if ((rc = batch_upload(*b))) {
        b->cases.pop_back();
        return rc;
    }
    return LBM355_OK;
}

int lbm355_batch_remove(lbm355_batch* b, lbm355_case* c)
{
    if (!b || !c) return fail(LBM355_EINVAL, "null argument");
    auto it = std::find(b->cases.begin(), b->cases.end(), c);
    if (it == b->cases.end()) return fail(LBM355_ENOTFOUND, "case is not a member of the batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->cases.erase(it);
    return batch_upload(*b);
}

}  // extern "C"

// several membership changes with one upload (the pool retires and admits
// cases between ticks)
int batch_update(lbm355_batch* b, const std::vector<lbm355_case*>& add, const std::vector<lbm355_case*>& remove)
{
    if (!b) return fail(LBM355_EINVAL, "null batch");
    if (add.empty() && remove.empty()) return LBM355_OK;
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (lbm355_case* c : remove) {
        auto it = std::find(b->cases.begin(), b->cases.end(), c);
        if (it == b->cases.end()) return fail(LBM355_ENOTFOUND, "case is not a member of the batch");
        b->cases.erase(it);
    }
    for (lbm355_case* c : add) {
        int rc;
        if ((rc = check_member(*b, c))) return rc;
        HIP_TRY(hipStreamSynchronize(c->stream));
        b->cases.push_back(c);
    }
    return batch_upload(*b);
}

extern "C" {

int lbm355_batch_size(const lbm355_batch* b, int* members, int64_t* bricks)
{
    if (!b) return fail(LBM355_EINVAL, "null batch");
    if (members) *members = int(b->cases.size());
    if (bricks) *bricks = b->n_list;
    return LBM355_OK;
}

int lbm355_batch_set_shape(lbm355_batch* b, int bpg)
{
    if (!b || !(bpg == 1 || bpg == 2 || bpg == 4 || bpg == 8)) return fail(LBM355_EINVAL, "bricks per group");
    b->bricks_per_group = bpg;
    drop_graph(*b);
    return LBM355_OK;
}

int lbm355_batch_set_per_wave(lbm355_batch* b, int per_wave)
{
    if (!b || per_wave < 1 || per_wave > 64) return fail(LBM355_EINVAL, "bricks per wavefront is 1..64");
    b->per_wave = per_wave;
    drop_graph(*b);
    return LBM355_OK;
}

int lbm355_batch_set_group_bytes(lbm355_batch* b, int64_t bytes)
{
    if (!b || bytes < 0) return fail(LBM355_EINVAL, "group bytes must be >= 0");
    b->group_bytes = bytes;
    regroup(*b);
    drop_graph(*b);
    return LBM355_OK;
}

int lbm355_batch_set_occupancy(lbm355_batch* b, int waves_per_cu)
{
    if (!b || !(waves_per_cu == 0 || (waves_per_cu >= 4 && waves_per_cu <= 32)))
        return fail(LBM355_EINVAL, "wavefronts per CU is 0 (no cap) or 4..32");
    b->waves_cap = waves_per_cu;
    drop_graph(*b);
    return LBM355_OK;
}

int lbm355_batch_config(const lbm355_batch* b, int out[8])
{
    if (!b || !out) return fail(LBM355_EINVAL, "null argument");
    out[0] = b->bricks_per_group;
    out[1] = b->nontemporal;
    out[2] = b->waves_cap;
    out[3] = int(b->group_bytes >> 20);
    out[4] = b->per_wave;
    out[5] = int(b->groups.size());
    out[6] = int(b->uploads);
    out[7] = int(b->captures);
    return LBM355_OK;
}

}  // extern "C"

// the library's error channel for the other host sources
int lbm355_set_error(int code, const std::string& msg) { return fail(code, msg); }
// device memory a step of the case sweeps, for the pool's scheduling
int64_t lbm355_case_step_bytes(const lbm355_case* c) { return c ? case_step_bytes(*c) : 0; }

// queue n steps of a batch on its stream without waiting (the pool enqueues
// every batch of a tick, then finishes them)
int batch_enqueue(lbm355_batch* b, int64_t n)
{
    if (!b) return fail(LBM355_EINVAL, "null batch");
    if (n <= 0 || b->cases.empty()) return LBM355_OK;
    // the batch's steps follow whatever is queued on each member's own stream
    // (steps, statistics resets, inflow updates): the streams are
    // independent, so without this wait both could touch one case at once
    for (lbm355_case* c : b->cases) {
        HIP_TRY(hipEventRecord(c->joined, c->stream));
        HIP_TRY(hipStreamWaitEvent(b->stream, c->joined, 0));
    }
    for (size_t i = 0; i < b->cases.size(); ++i) {
        CaseDev* dc = static_cast<CaseDev*>(b->devcases.p) + i;
        // a case stepped on its own since the last batch call: re-anchor it
        const int64_t want = b->launches - b->cases[i]->steps;
        if (want != b->start[i]) {
            LAUNCH_TRY(launch_set_u64(&dc->start, uint64_t(want), b->stream));
            b->start[i] = want;
        }
        if (b->epoch[i] != b->cases[i]->stats_first) {
            LAUNCH_TRY(launch_set_u64(&dc->stats_epoch, uint64_t(b->cases[i]->stats_first), b->stream));
            b->epoch[i] = b->cases[i]->stats_first;
        }
        // the per-step inflow hook (lbm355_case_set_inflow, updateKernelVelocities
        // of sim_tcpc.cu:657-663) writes the case's own copy of its constants
        const CaseDev& h = b->cases[i]->host;
        if (h.inflow_xm != b->in_xm[i]) {
            uint64_t v;
            std::memcpy(&v, &h.inflow_xm, 8);
            LAUNCH_TRY(launch_set_u64(&dc->inflow_xm, v, b->stream));
            b->in_xm[i] = h.inflow_xm;
        }
        if (h.inflow_xp != b->in_xp[i]) {
            uint64_t v;
            std::memcpy(&v, &h.inflow_xp, 8);
            LAUNCH_TRY(launch_set_u64(&dc->inflow_xp, v, b->stream));
            b->in_xp[i] = h.inflow_xp;
        }
        // rho,u of every case at the last step of this call
        LAUNCH_TRY(launch_set_u64(&dc->macro_step, uint64_t(b->cases[i]->steps + n - 1), b->stream));
    }
    // steps per group per replay: long enough that refilling the Infinity
    // Cache at a group switch is a few percent of the chunk
    const int chunk = 64;
    const int grid = std::max(batch_grid(*b), b->graph_grid);
    int64_t done = 0;
    if (b->use_graph && n >= chunk) {
        if (!b->exec) {
            // one graph of `chunk` consecutive launches (launch index base +
            // 0..chunk-1) of the group whose range the descriptor holds;
            // replayed once per group, then the launch counter advances
            StepLaunch L = batch_step_launch(*b, grid);
            std::lock_guard<std::mutex> capture_lock(g_device_mu);
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int r = 0;
            for (int s = 0; s < chunk && r == 0; ++s) {
                L.rel = s;
                r = launch_step(b->storage, b->model, L, b->stream);
            }
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (r != 0 || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(LBM355_EDEVICE, r ? kernel_error() : hipGetErrorString(e));
            }
            b->graph = g;
            HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->graph_grid = grid;
            ++b->captures;
        }
        while (n - done >= chunk) {
            int rc;
            for (const auto& G : b->groups) {
                if ((rc = set_range(*b, G.begin, G.end))) return rc;
                HIP_TRY(hipGraphLaunch(b->exec, b->stream));
            }
            LAUNCH_TRY(launch_advance(static_cast<int64_t*>(b->base.p), chunk, b->stream));
            done += chunk;
            b->launches += chunk;
        }
    }
    // the remaining steps group by group; every group replays the same
    // launch indices, so each member sees its own consecutive steps
    if (n > done) {
        StepLaunch L = batch_step_launch(*b, grid);
        int rc;
        for (const auto& G : b->groups) {
            if ((rc = set_range(*b, G.begin, G.end))) return rc;
            for (int64_t s = 0; s < n - done; ++s) {
                L.rel = s;
                LAUNCH_TRY(launch_step(b->storage, b->model, L, b->stream));
            }
        }
        LAUNCH_TRY(launch_advance(static_cast<int64_t*>(b->base.p), n - done, b->stream));
        b->launches += n - done;
    }
    // and whatever is queued for a member afterwards follows the batch's steps
    HIP_TRY(hipEventRecord(b->done, b->stream));
    for (lbm355_case* c : b->cases) HIP_TRY(hipStreamWaitEvent(c->stream, b->done, 0));
    for (lbm355_case* c : b->cases) {
        if (c->d.stats_every > 0) {
            const int64_t k = c->d.stats_every;
            c->stats_steps += (c->steps + n) / k - c->steps / k;
        }
        c->steps += n;
        c->macro_valid_step = -2;  // rho,u of the last step land once the stream is done
        c->strain_valid = false;
    }
    return LBM355_OK;
}

// wait for a batch's queued steps; the macro fields of every member then hold
// rho,u of its last step
int batch_finish(lbm355_batch* b)
{
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (lbm355_case* c : b->cases) c->macro_valid_step = c->steps - 1;
    return LBM355_OK;
}

extern "C" {

int lbm355_batch_step(lbm355_batch* b, int64_t n)
{
    int rc = batch_enqueue(b, n);
    if (rc) return rc;
    return n > 0 ? batch_finish(b) : LBM355_OK;
}

namespace {
// wall time of `steps` batch steps after 2 warm-up steps (which also capture
// the graph), ms
int time_batch(lbm355_batch* b, int steps, hipEvent_t e0, hipEvent_t e1, float* ms)
{
    int rc;
    if ((rc = batch_enqueue(b, 2)) || (rc = batch_finish(b))) return rc;
    if (b->use_graph && steps >= 64 && (rc = batch_enqueue(b, 64))) return rc;  // graph capture outside the window
    if ((rc = batch_finish(b))) return rc;
    HIP_TRY(hipEventRecord(e0, b->stream));
    if ((rc = batch_enqueue(b, steps))) return rc;
    HIP_TRY(hipEventRecord(e1, b->stream));
    if ((rc = batch_finish(b))) return rc;
    HIP_TRY(hipEventElapsedTime(ms, e0, e1));
    return LBM355_OK;
}
}  // namespace

int lbm355_batch_autotune(lbm355_batch* b, int* best)
{
    if (!b) return fail(LBM355_EINVAL, "null batch");
    if (b->cases.empty()) return fail(LBM355_EINVAL, "empty batch");
    // real steps of the batch, a whole number of statistics groups per window
    // so both AA parities and the same number of flushes count; coordinate
    // search: launch shape x cache policy x wavefront cap, then bricks per
    // wavefront (the persistent loop), then the group size
    int steps = 8;
    for (const lbm355_case* c : b->cases) steps = std::max(steps, tune_window(c->d.stats_every));
    int rc = batch_enqueue(b, kTuneWarmup);
    if (rc || (rc = batch_finish(b))) return rc;
    EventSet es(2);
    HIP_TRY(hipEventCreate(&es.ev[0]));
    HIP_TRY(hipEventCreate(&es.ev[1]));
    float best_ms = 0.f;
    int best_bpg = b->bricks_per_group, best_nt = b->nontemporal, best_cap = b->waves_cap, best_pw = 1;
    bool first = true;
    b->per_wave = 1;
    for (int k = 0; k < kTuneConfigs; ++k) {
        const TuneConfig t = tune_config(k);
        b->bricks_per_group = t.bpg;
        b->nontemporal = t.nt;
        b->waves_cap = t.cap;
        drop_graph(*b);
        float ms = 0.f;
        if ((rc = time_batch(b, steps, es.ev[0], es.ev[1], &ms))) return rc;
        if (first || ms < best_ms) { best_ms = ms; best_bpg = t.bpg; best_nt = t.nt; best_cap = t.cap; }
        first = false;
    }
    b->bricks_per_group = best_bpg;
    b->nontemporal = best_nt;
    b->waves_cap = best_cap;
    for (int pw : {2, 4}) {
        b->per_wave = pw;
        drop_graph(*b);
        float ms = 0.f;
        if ((rc = time_batch(b, steps, es.ev[0], es.ev[1], &ms))) return rc;
        if (ms < best_ms) { best_ms = ms; best_pw = pw; }
    }
    b->per_wave = best_pw;
    const int64_t sizes[3] = {0, int64_t(256) << 20, int64_t(128) << 20};
    int64_t best_group = b->group_bytes;
    bool gfirst = true;
    for (int64_t gb : sizes) {
        b->group_bytes = gb;
        regroup(*b);
        drop_graph(*b);
        float ms = 0.f;
        if ((rc = time_batch(b, 2 * steps, es.ev[0], es.ev[1], &ms))) return rc;
        if (gfirst || ms < 2.f * best_ms) { best_ms = ms * 0.5f; best_group = gb; }
        gfirst = false;
    }
    b->group_bytes = best_group;
    regroup(*b);
    drop_graph(*b);
    if (best) *best = best_bpg | (best_nt << 8) | (best_cap << 16) | int(b->group_bytes >> 24) << 24;
    return LBM355_OK;
}

int lbm355_case_autotune(lbm355_case* c, int* best)
{
    if (!c) return fail(LBM355_EINVAL, "null case");
    // kTuneWarmup steps first; stage 1: every configuration on 2 warm-up steps
    // and a timed window of W back-to-back steps (tune_window); stage 2: the
    // kTuneFinalists fastest, kTunePasses interleaved passes timing 4W steps,
    // each keeping its fastest pass.  Single windows of a ~20 us step are within launch jitter and
    // clock drift of each other: with 3 finalists and two passes of 2W steps
    // the 256x192x128 junction settled on 4 bricks per workgroup and a cap of
    // 16 in one bench run (19.45 us per step) against 18.4 us for 1 brick.
    const int W = tune_window(c->d.stats_every);
    std::vector<std::pair<double, int>> stage1;
    int rc;
    if ((rc = lbm355_case_step(c, kTuneWarmup))) return rc;
    for (int k = 0; k < kTuneConfigs; ++k) {
        const TuneConfig t = tune_config(k);
        c->bricks_per_group = t.bpg;
        c->nontemporal = t.nt;
        c->waves_cap = t.cap;
        if ((rc = lbm355_case_step(c, 2))) return rc;  // warm-up
        double step_ms = 0.0;
        // time_kernel(2n) times its first n steps back to back: W steps, one
        // whole statistics group.  Timing only W/2 put the group's flush step
        // (about twice a plain step) into the window of every other
        // configuration; with 18 steps per configuration that phase repeats
        // every 8, so the same four shapes carried the flush in every cap
        // pass and lost stage 1 (512x384x256 KBC: the three fastest shapes,
        // 145-148 us per step, missed the finalists and 157 us was kept)
        if ((rc = lbm355_case_time_kernel(c, 2 * W, &step_ms, nullptr))) return rc;
        stage1.push_back({step_ms, k});
    }
    std::sort(stage1.begin(), stage1.end());
    const int nf = std::min(kTuneFinalists, int(stage1.size()));
    std::vector<double> fin(size_t(nf), 0.0);
    for (int pass = 0; pass < kTunePasses; ++pass)
        for (int r = 0; r < nf; ++r) {
            const TuneConfig t = tune_config(stage1[size_t(r)].second);
            c->bricks_per_group = t.bpg;
            c->nontemporal = t.nt;
            c->waves_cap = t.cap;
            if ((rc = lbm355_case_step(c, 2))) return rc;
            double step_ms = 0.0;
            if ((rc = lbm355_case_time_kernel(c, 8 * W, &step_ms, nullptr))) return rc;  // 4W back to back
            if (pass == 0 || step_ms < fin[size_t(r)]) fin[size_t(r)] = step_ms;
        }
    int best_r = 0;
    for (int r = 1; r < nf; ++r)
        if (fin[size_t(r)] < fin[size_t(best_r)]) best_r = r;
    const TuneConfig bt = tune_config(stage1[size_t(best_r)].second);
    const int best_bpg = bt.bpg, best_nt = bt.nt, best_cap = bt.cap;
    c->bricks_per_group = best_bpg;
    c->nontemporal = best_nt;
    c->waves_cap = best_cap;
    if (best) *best = best_bpg | (best_nt << 8) | (best_cap << 16);
    return LBM355_OK;
}

int lbm355_batch_destroy(lbm355_batch* b)
{
    delete b;
    return LBM355_OK;
}

}  // extern "C"
