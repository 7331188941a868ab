// imh_host.h -- host-side structures shared by the C ABI implementation files.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "imh_common.h"
#include "imh_kernels.h"

namespace imh {

struct FluidTile {
    int bs, kmax, PX, TL, K, xmajor;
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
    long long eoff[imh::NOBJ][2][3];
    long long* eoff_dev = nullptr;
    imh::SimDev* sim = nullptr;
    imh::SimParams* par = nullptr;
    imh::SimParams* par_uncond = nullptr;
    imh::SimParams host_par;
    double* csum_part = nullptr;
    double* csum_out = nullptr;
    bool obj_init[imh::NOBJ] = {};
    bool fluid_init = false;
    bool have_state = false;
    bool mhd_kernel = false;
    bool no_graphs = false;
    int mag_is_zero[3] = {1, 1, 1};
    double fade[imh::MAX_FADE] = {};
    double* fade_dev = nullptr;
    int host_iteration = 1;
    imh::FluidTile ftile[3];
    imh::MagTile mtile[3];
    hipGraphExec_t graph[6] = {};
};

// internal entry points (force = run regardless of the loop's running flag)
int imh_fluid_sweep_internal(imh_ctx* c, int X, int force);
int imh_magnet_sweep_internal(imh_ctx* c, int X, const int I[2], int force);
int imh_time_update_internal(imh_ctx* c, int step);
int imh_flux_internal(imh_ctx* c, int order, int force);
