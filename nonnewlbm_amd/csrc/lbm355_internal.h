// lbm355_internal.h - shared definitions of the device kernels and the host
// driver of liblbm355.
//
// Data layout in HBM (per case)
//   The grid is cut into kBX x kBY x kBZ bricks (8 x 8 x 1); one brick = 64
//   cells = one wavefront.
//   Bricks are stored only where the AA update touches memory (the non-wall
//   cells and their 26-neighbourhood); active bricks (holding at least one
//   non-wall cell) come first, z slab by z slab and row-major in a slab, then
//   the storage-only shell bricks, then two bricks that stand in for every
//   brick that is not stored: the "dummy" (index n_stored), read by boundary
//   partner lookups of absent cells and never written, and the "sink"
//   (n_stored + 1), the target of the neighbour table for absent bricks.  The
//   odd step's writes from boundary-plane cells that leave the grid land in
//   the sink or in outside lanes of stored bricks, which nothing reads.
//   populations : f[brick][q][lane], lane = i + kBX (j + kBY k), one 64-element
//                 run per (brick, q): every wavefront load is one contiguous
//                 64 x sizeof(T) segment (256 B at FP32).
//   AA pattern  : one buffer.  An even step reads and writes the cell's own 27
//                 slots (post-collision f_q goes to slot opp(q)); an odd step
//                 reads slot opp(q) of the upstream neighbour and writes slot q
//                 of the downstream neighbour.  Full-way bounce-back wall nodes
//                 (solver/lbm3d/lbm_streaming.h:78-107) are exactly a no-op in
//                 this scheme, so wall cells need storage but no thread.
#pragma once

#include <cstdint>

namespace lbm355 {

constexpr int kLanes = 64;           // cells per brick = wavefront width
constexpr int kSpareBricks = 2;      // dummy + sink after the stored bricks
// brick shape.  One cell in z: the odd AA step's c_z shifts then move whole
// population runs between bricks instead of splitting their lines, and a
// wavefront's shifted accesses touch fewer partial lines than with 4x4x4
// bricks (measured at 512x384x256: 161 us per FP32 step against 181 us;
// 16x4x1 bricks give 165 us).  Each axis that carries BC faces (x, y) needs
// at least 3 cells for the brick origin shift; the z faces are always walls.
constexpr int kBX = 8, kBY = 8, kBZ = 1;
constexpr int kSX = 3, kSY = 3;      // log2 kBX, log2 kBY
static_assert(kBX * kBY * kBZ == kLanes && (1 << kSX) == kBX && (1 << kSY) == kBY, "brick shape");
static_assert(kBX >= 3 && kBY >= 3, "x and y carry BC faces");

// lane of brick-local cell (i, j, k), wrapped into the brick
__host__ __device__ constexpr int brick_lane(int i, int j, int k)
{
    return (i & (kBX - 1)) | ((j & (kBY - 1)) << kSX) | ((k & (kBZ - 1)) << (kSX + kSY));
}
constexpr int kQ = 27;               // D3Q27
constexpr int kStatFields = 6;       // sum rho u (3), sum (rho u)^2 (3)
constexpr int kMacroFields = 4;      // rho, ux, uy, uz
constexpr int kStatBlock = 16;       // samples per addition to the FP64 sums (non-FP64 storage)
constexpr int kRingFields = 3;       // rho u of one sample in a statistics ring slot

// cell labels (the reference map values, sim_tcpc.cu:194-203, plus 7/8 for
// outflow planes on x faces) and 255 = outside the grid
enum : uint8_t {
    L_FLUID = 0, L_WALL = 1, L_IN_XM = 3, L_IN_XP = 4, L_OUT_YM = 5, L_OUT_YP = 6,
    L_OUT_XP = 7, L_OUT_XM = 8, L_OUTSIDE = 255
};

__host__ __device__ constexpr bool is_bc_label(int l) { return l >= 3 && l <= 8; }
__host__ __device__ constexpr bool is_inflow_label(int l) { return l == L_IN_XM || l == L_IN_XP; }

// per-case constants and pointers read by the kernels (device resident so a
// captured graph can be replayed while the host changes inflow or output
// requests between replays)
struct CaseDev {
    void* f;                  // populations, storage type of the launch
    double* stats;            // [active brick][6][64] FP64 statistics sums, or null
    float* stats_part;        // [kStatBlock - 1][active brick][3][64] FP32 rho u samples of the open group
    void* macro;              // [active brick][4][64] rho,u (compute type), or null
    void* strain;             // [active brick][6][64] strain (compute type), or null
    void* force;              // [active brick][3][64] body force (fd-force), or null
    const int32_t* nbr;       // [active brick][27] element offset (brick * 27 * 64) of the neighbour bricks
    const uint8_t* label;     // [stored brick + dummy + sink][64]
    const int32_t* bc_base;   // [active brick] first BC slot of the brick or -1
    const int32_t* bc_cell;   // [nbc] active brick * 64 + lane of each BC slot
    const int32_t* bc_read;   // [nbc][2][27] element offsets of the partner reads
    const double* bc_factor;  // [nbc] parabolic profile factor, or the inlet area (plug)
    const uint8_t* bc_label;  // [nbc]
    const int16_t* origin;    // [active brick][3] brick coordinates (cells / 4)
    int64_t start;            // launch index of the case's step 0: step = launch - start
    int64_t f_bytes;          // size of the population buffer (< 4 GiB)
    int64_t stats_bytes;      // size of the statistics buffer (< 4 GiB)
    int64_t stats_epoch;      // step index of the last statistics reset
    int32_t n_active;         // active bricks
    int32_t n_stored;         // stored bricks (dummy brick = n_stored, sink = n_stored + 1)
    int32_t n_bc;
    int32_t nx, ny, nz;
    double nu_ref;            // lattice viscosity of the collision
    double cy_nu0, cy_lambda, cy_a, cy_n;
    // wave-uniform constants of the neq Carreau-Yasuda solve, computed once at
    // case creation: the exponent (n - 1) / a, log2 lambda and log2 of the
    // larger / smaller of the relaxation times 3 nu_ref + 1/2 and 3 nu0 + 1/2
    double cy_e, cy_log2_lambda, cy_log2_tau_max, cy_log2_tau_min;
    // the same constants rounded once to FP32 for the FP32 kernels
    float nu_ref_f, cy_nu0_f, cy_a_f, cy_n_f, cy_e_f, cy_log2_lambda_f, cy_log2_tau_max_f, cy_log2_tau_min_f;
    double inflow_xm, inflow_xp;   // IVC / SVC scale: flux (plug) or u_max (parabolic)
    double pulse_amp, pulse_period;
    int32_t cy_mode;          // 0 off, 1 neq, 2 fd-force (with cy_enabled)
    int32_t stats_every;      // 0 = off
    int64_t macro_step;       // write rho,u at this step index (-1 = never)
    int32_t macro_every;      // also every k steps (0 = never)
    int32_t plug;             // 1: u = inflow / area (sim_tcpc.cu:298,306)
};

struct BrickRef {
    uint32_t case_id;
    uint32_t brick;
};

// device working set above which the even step streams its populations with
// the non-temporal hint (measured crossover on MI355X between 331 MB, where the
// Infinity Cache still serves the next step, and 509 MB)
constexpr int64_t kNontemporalBytes = int64_t(448) << 20;
// batches step their members in groups of at most this much device memory
constexpr int64_t kBatchGroupBytes = int64_t(256) << 20;

// device-resident descriptor of a batch: its concatenated brick list, its
// members' constants, and the brick range of the group the next persistent
// launches step (changed between launches or graph replays, never inside)
struct BatchDev {
    const BrickRef* list;
    CaseDev* cases;
    int32_t begin, end;
};

// launch descriptor of one fused step over a brick list.  With batch == null
// it is one brick per wavefront over list[0, n_list) (k_step); otherwise the
// persistent form over the batch descriptor's range with a grid of n_list
// workgroups (k_step_loop)
struct StepLaunch {
    const BrickRef* list;
    int32_t n_list;
    CaseDev* cases;           // device array
    int32_t n_cases;
    const int64_t* base;      // device launch counter; this launch is *base + rel
    int64_t rel;
    int32_t bricks_per_group; // 1, 2, 4, 8 wavefronts per workgroup
    int32_t nontemporal;      // 1: even-step population accesses with the nt hint
    int32_t lds_bytes;        // dynamic LDS per workgroup (unused by the kernel): caps wavefronts per CU
    const BatchDev* batch;    // batch launch (device descriptor), or null
    int32_t per_wave;         // batch launches: 1 one brick per wavefront (k_step_range), k > 1 the
                              // persistent loop (k_step_loop, grid sized for about k bricks per wavefront)
};

// storage kinds (same values as LBM355_F64 ...)
enum Storage : int { S_F64 = 0, S_F32 = 1, S_F16 = 2, S_BF16 = 3 };
// the two-byte storage types use the balanced AA addressing: pair (cell P,
// slot s) lives at cell P - c_s for s < 13 and at P otherwise, so that both
// AA steps shift half of the directions (see balanced_loads)
template <int S>
constexpr bool kBalanced = S == S_F16 || S == S_BF16;
constexpr bool balanced_storage(int s) { return s == S_F16 || s == S_BF16; }
enum Model : int { M_CUM = 0, M_BGK = 1, M_MRT = 2, M_KBC = 3 };

// D3Q27 lattice weights w_q, q = 9(cx+1)+3(cy+1)+(cz+1)
__host__ __device__ constexpr double lattice_weight(int q)
{
    return ((q / 9 - 1) * (q / 9 - 1) + ((q / 3) % 3 - 1) * ((q / 3) % 3 - 1) + (q % 3 - 1) * (q % 3 - 1)) == 0
               ? 8.0 / 27.0
           : ((q / 9 - 1) * (q / 9 - 1) + ((q / 3) % 3 - 1) * ((q / 3) % 3 - 1) + (q % 3 - 1) * (q % 3 - 1)) == 1
               ? 2.0 / 27.0
           : ((q / 9 - 1) * (q / 9 - 1) + ((q / 3) % 3 - 1) * ((q / 3) % 3 - 1) + (q % 3 - 1) * (q % 3 - 1)) == 2
               ? 1.0 / 54.0
               : 1.0 / 216.0;
}

// FP16-offset storage keeps (f_q - w_q) * 2^e_q in IEEE half, e_q chosen per
// velocity class so that w_q * 2^e_q = 16/27 for every q: the stored value is
// (f_q / w_q - 1) * 16/27, the same scale for all 27 directions, and small
// deviations of the 1/216 corner weights stay out of the half subnormals
__host__ __device__ constexpr int half_scale_exp(int q)
{
    return ((q / 9 - 1) * (q / 9 - 1) + ((q / 3) % 3 - 1) * ((q / 3) % 3 - 1) + (q % 3 - 1) * (q % 3 - 1)) == 0 ? 1
           : ((q / 9 - 1) * (q / 9 - 1) + ((q / 3) % 3 - 1) * ((q / 3) % 3 - 1) + (q % 3 - 1) * (q % 3 - 1)) == 1 ? 3
           : ((q / 9 - 1) * (q / 9 - 1) + ((q / 3) % 3 - 1) * ((q / 3) % 3 - 1) + (q % 3 - 1) * (q % 3 - 1)) == 2 ? 5
                                                                                                              : 7;
}

// launchers (lbm355_kernels.hip); all asynchronous on `stream`
// one fused step over the launch's brick list
int launch_step(int storage, int model, const StepLaunch& L, void* stream);
int launch_copy(void* dst, const void* src, int64_t bytes, void* stream);
int launch_advance(int64_t* base, int64_t k, void* stream);
int launch_fill_populations(int storage, void* f, int64_t n_bricks, void* stream);
int launch_strain(int storage, const CaseDev* dev_case, int n_active, void* stream);
// fd-force Carreau-Yasuda mode: velocity pre-pass, strain and body force of
// the step L is about to run (launch them right before it on the same stream)
int launch_fd_force_prepass(int storage, const StepLaunch& L, const CaseDev* dev_case, int n_active, void* stream);
// plane fluxes of the BC cells into dev_out[4] (accumulated: clear it first);
// mean = 1 sums the statistics (FP64 sums + `open` ring samples) instead of u
int launch_flux(int storage, const CaseDev* dev_case, int n_bc, double* dev_out, int mean, int open, void* stream);
// difference norms of two cases with the same brick index into
// partials[4 * compare_blocks(n_active) + 0..3] (the partials before them are
// scratch of the same call): sum |u_a - u_b|^2, |u_b|^2, (rho_a - rho_b)^2,
// (rho_b - 1)^2 over the non-wall cells
int compare_blocks(int n_active);
int launch_compare(int storage_a, int storage_b, const CaseDev* a, const CaseDev* b, int n_active, double* partials,
                   void* stream);
// device address of the static compare scratch on `device` (looked up once
// per device) and the most workgroups of partials it holds
int compare_scratch(int device, double** ptr, int* max_blocks);
int launch_set_u64(void* dev_ptr, uint64_t value, void* stream);
int launch_set_batch(BatchDev* dev_ptr, const BatchDev& value, void* stream);
int launch_zero(void* ptr, int64_t bytes, void* stream);
const char* kernel_error();

}  // namespace lbm355
