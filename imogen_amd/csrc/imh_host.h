// imh_host.h -- host-side structures shared by the C ABI implementation files.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "imh_common.h"
#include "imh_kernels.h"
#include "imogen_hip_ext.h"

namespace imh {

// Kernel timing with HIP events recorded on the launching stream around every launch,
// with the launch's algorithmic HBM bytes (compulsory reads and writes of the fields).
struct ProfRec {
    int kind;
    double bytes;
    hipEvent_t a, b;
};
struct Profiler {
    std::vector<ProfRec> recs;
    bool failed = false;
};
void prof_begin(Profiler* p, hipStream_t st, int kind, double bytes);
void prof_end(Profiler* p, hipStream_t st);
// Waits for the events, sums them per kind into ms/bytes/count[IMH_NKIND] and frees them.
int prof_collect(Profiler* p, double* ms, double* bytes, int* count);

struct FluidTile {
    int bs, kmax, PX, TL, K, xmajor;
    int column;  // sweep by the column march (fluid_column.hip) when the launch is lean
};

struct MagTile {
    int PX, TL;
};

struct TileOverride {
    int valid, bs, kmax, px;
};

extern thread_local std::string g_last_error;
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
int validate_config(const imh_config* c);
void sweep_order(int iteration, int order, int perms_order, int dirs[3], int mi[3][2]);
FluidTile choose_fluid_tile(int N, int d, int esz, const TileOverride* ov);
bool fluid_variant_exists(int esz, int bs, int kmax);
MagTile choose_mag_tile(int N);
int column_possible(const imh_config* cfg, int d);

}  // namespace imh

struct imh_ctx {
    imh_config cfg;
    int n[3];
    int ndims;
    long long ncell;
    long long nface[3];
    int active_lin[3];
    hipStream_t st = nullptr;
    void* f[IMH_NFIELDS] = {};
    void* scratch = nullptr;        // mag(I).flux(X) for the CT step
    double* edges = nullptr;
    size_t edges_len = 0;
    size_t edges_cap = 0;               // doubles allocated at edges (>= edges_len)
    // dimensions whose pencils exceed the fused sweep's tile: swept by fluid_long.hip,
    // with its scratch area (fluid_long_bytes of the longest such sweep)
    bool long_dim[3] = {false, false, false};
    void* long_scratch = nullptr;
    size_t long_bytes = 0;
    long long eoff[imh::NOBJ][2][3];
    long long* eoff_dev = nullptr;
    imh::SimDev* sim = nullptr;
    imh::SimParams* par = nullptr;
    imh::SimDev** sim_ptr = nullptr;        // device array {sim}
    imh::SimParams** par_ptr = nullptr;     // device array {par}
    imh::SimParams host_par;
    double* csum_part = nullptr;
    double* csum_out = nullptr;
    bool obj_init[imh::NOBJ] = {};
    bool fluid_init = false;
    bool squeezed = false;   // the grid was relabelled by squeeze_grid (singleton x or y)
    int sweep_kind = 0;             // IMH_SWEEP_GENERAL / _WAVE / _WAVE_FAST
    int num_cu = 256;               // compute units of the device
    size_t max_lds = 64 * 1024;     // LDS bytes per workgroup of the device
    bool have_state = false;
    bool mhd_kernel = false;
    bool no_graphs = false;
    bool external_stream = false;   // the stream belongs to a batch
    int mag_is_zero[3] = {1, 1, 1};
    double fade[imh::MAX_FADE] = {};
    double* fade_dev = nullptr;
    int host_iteration = 1;
    imh::FluidTile ftile[3];
    imh::MagTile mtile[3];
    hipGraphExec_t graph[6] = {};
    // Kernel argument blocks live in device memory (the kernels take a pointer, so one
    // kernel serves a single simulation and a batch).  Eager launches stage them through
    // a pinned ring copied on the stream; captured launches take a slot of the row's
    // graph buffer, filled once after the capture.
    char* arg_host = nullptr;           // pinned, ARG_RING slots
    char* arg_dev = nullptr;            // device, ARG_RING slots
    int arg_next = 0;
    int capture_row = -1;               // >= 0 while a row's graph is being captured
    int cap_count = 0;
    char* cap_host = nullptr;           // host image of the captured row's slots
    char* cap_dev[6] = {};              // device, CAP_SLOTS slots per row
    imh::Profiler* prof = nullptr;      // set while imh_profile_iterations runs
    // source terms (source.m): fixed gravitational potential and radiation
    double* grav = nullptr;             // grav.array (FP64), nullptr: gravity off
    // non-uniform DGRID{d}: one line of n[d] spacings on the device, nullptr: uniform
    double* dline_dev[3] = {nullptr, nullptr, nullptr};
    double* grav_ghost = nullptr;       // 6 ghost values per pencil end (source.hip)
    double mass_threshold = 0.0;        // run.fluid.MASS_THRESHOLD
    int rad_type = 0;                   // 0 'empty', 1 'optically_thin'
    double rad_exponent = 0.5;
    double rad_strength = 0.0;
    // run.fluid.viscosity (ArtificialViscosity.m): 0 'empty', 1 'neumann_richtmyer'
    int visc_type = 0;
    double visc_lin = 0.0, visc_quad = 0.0;
};
// The fluid sweep along d takes the multi-launch form (fluid_long.hip): pencils longer
// than a workgroup holds, or the Caramana-Shashkov-Whalen viscosity with a fade end.
inline bool long_sweep(const imh_ctx* c, int d) {
    return c->long_dim[d] ||
           (c->visc_type == 2 && (c->cfg.bc[0][d] == IMH_BC_FADE || c->cfg.bc[1][d] == IMH_BC_FADE));
}


namespace imh {
constexpr size_t ARG_SLOT = 512;       // bytes per staged argument block
constexpr int ARG_RING = 256;
constexpr int CAP_SLOTS = 96;          // launches in one captured iteration (44 at most)
}  // namespace imh

// Device copy of an argument block for the next launch on c->st (nullptr on failure,
// with the error recorded).
const void* imh_stage_args(imh_ctx* c, const void* a, size_t n);

// Parameters of a device-generated initial state (initial.hip).
struct IcParams {
    int kind;
    int n[3];
    int ip[4];            // integer parameters (split cells, counts)
    double dx[3];
    double gamma;
    double p[24];         // real parameters
    unsigned long long seed;
};
namespace imh {
hipError_t launch_ic_fill(const IcParams& P, void* const f[8], bool fp64, hipStream_t st);
hipError_t launch_ic_edges(const IcParams& P, int field, int side, int d, double* edges,
                           long long off, long long nface, hipStream_t st);
}  // namespace imh

// Arguments of the source step (source.hip); plain values, passed by value.
struct SourceArgs {
    const double* phi;        // potential, nullptr: no gravity
    double* ghosts;           // [sum of nface over d < ndims][6]
    double* edges;            // the context's edge stores
    long long eoff[2][3];     // offsets of the potential's trackers, -1 where none
    long long goff[3];        // first ghost face of each dim
    long long nface[3];
    int n[3];
    int ndims;
    int modes[2][3];
    int active_lin[3];
    int bc_inf;
    int rad_type;
    const imh::SimDev* sim;
    const double* fade;       // fade_shift.m coefficients (device), bcInfinity long
    double mass_threshold;
    double dgrid[3];
    const double* dxl[3];     // non-uniform spacing lines (device), nullptr: dgrid[d]
    double gamma;
    double rad_strength;
    double rad_exponent;
    imh::MagUse use;
};
namespace imh {
hipError_t launch_source(const SourceArgs& a, void* const f[IMH_NFIELDS], bool fp64,
                         hipStream_t st);
// edges[f] = field at the lower (side 0) or upper (side 1) end of pencil f along d
hipError_t launch_edges_from_field(const void* field, bool fp64, const int n[3], int d,
                                   int side, double* edges, long long nface, hipStream_t st);
}  // namespace imh

// internal entry points (force = run regardless of the loop's running flag)
int imh_fluid_sweep_internal(imh_ctx* c, int X, int force);
int imh_magnet_sweep_internal(imh_ctx* c, int X, const int I[2], int force);
int imh_time_update_internal(imh_ctx* c, int step);
int imh_flux_internal(imh_ctx* c, int order, int force);

// argument blocks of one simulation (imh_api.cpp), shared with the batched path
template <typename T>
imh::FluidArgs<T> make_fluid_args(const imh_ctx* c, int d, int force);
template <typename T>
imh::MagnetArgs<T> make_magnet_args(const imh_ctx* c, int dX, int dI, int force);
template <typename T>
imh::CellMagArgs<T> make_cellmag_args(const imh_ctx* c, int k, int force);
template <typename T>
imh::CflArgs<T> make_cfl_args(const imh_ctx* c);
int imh_iteration_body(imh_ctx* c, int conditional);
int imh_source_internal(imh_ctx* c);
// The source step's kernels of one context on stream st (no check of what is active).
hipError_t imh_launch_source(imh_ctx* c, hipStream_t st);
double source_bytes(const imh_ctx* c);
void imh_ctx_free(imh_ctx* c);
// A new context for cfg on the device memory of a finished one (same cell count,
// precision, field set, room for the edge stores); see imh_api.cpp.
int imh_ctx_recycle(imh_ctx* old, const imh_config* cfg, imh_ctx** out);
