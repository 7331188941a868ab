// imogen_mex.cpp -- MATLAB MEX gateway onto the C ABI (include/imogen_hip.h).
//
// Built only where MATLAB's mex.h is present (imogen_amd/mex/build_mex.sh); neither
// MATLAB nor mex.h exists in this image.  It follows the reference's MEX convention
// (gravity/utils/cudaCommon.cu:20-56, cudaTVDStep.cu:33): one mexFunction, arguments as
// mxArrays, results returned as new arrays or applied in place with nlhs = 0, errors
// raised with mexErrMsgIdAndTxt.  The device state stays resident between calls inside
// the library context whose handle MATLAB keeps (as a uint64 scalar) in the run object,
// the role the reference's static MEX state plays (freezeAndPtot.cu:17-18,34-40).
//
//   h = imogen_mex('create', cfg)                 cfg: struct with the imh_config fields
//   imogen_mex('upload', h, mass, mom, ener, mag) mom/mag: 3 x [nx ny nz] or cell arrays
//   imogen_mex('fluid', h, X)                     relaxingFluid(run,...,X)
//   imogen_mex('magnet', h, X, [I1 I2])           magnetFlux(run,...,X,I)
//   imogen_mex('time_update', h, i)               run.time.update(...,i)
//   imogen_mex('flux', h, order)                  flux(run,...,order)
//   imogen_mex('iteration', h)                    one imogen.m loop body
//   n = imogen_mex('run', h, maxIter)             the imogen.m loop
//   a = imogen_mex('download', h, fieldId, numel) FluidArray/MagnetArray .array (a column;
//                                                 reshape to the grid on the MATLAB side)
//   [t, dt, iter, gi] = imogen_mex('time', h)
//   imogen_mex('source', h)                       source(run,...)
//   imogen_mex('setGravity', h, phi, threshold)   fixed potential, MASS_THRESHOLD
//   imogen_mex('setRadiation', h, type, a, imax)  run.fluid.radiation
//   imogen_mex('setViscosity', h, type, c1, c2)   run.fluid.viscosity
//   imogen_mex('destroy', h)
#include <cstdint>
#include <cstring>
#include <string>

#include "imogen_hip.h"
#include "mex.h"

namespace {

void check(int rc) {
    if (rc != IMH_OK) mexErrMsgIdAndTxt("Imogen:hip", "%s (status %d)", imh_last_error(), rc);
}

imh_ctx* handle(const mxArray* a) {
    if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1)
        mexErrMsgIdAndTxt("Imogen:hip", "expected a context handle (uint64 scalar)");
    uint64_t v = *static_cast<const uint64_t*>(mxGetData(a));
    return reinterpret_cast<imh_ctx*>(static_cast<uintptr_t>(v));
}

double field_or(const mxArray* s, const char* name, double dflt) {
    const mxArray* f = mxGetField(s, 0, name);
    return f ? mxGetScalar(f) : dflt;
}

void field_vec(const mxArray* s, const char* name, double* out, int n) {
    const mxArray* f = mxGetField(s, 0, name);
    if (!f || (int)mxGetNumberOfElements(f) < n)
        mexErrMsgIdAndTxt("Imogen:hip", "cfg.%s needs %d elements", name, n);
    const double* p = mxGetPr(f);
    for (int i = 0; i < n; ++i) out[i] = p[i];
}

imh_config parse_config(const mxArray* s) {
    if (!mxIsStruct(s)) mexErrMsgIdAndTxt("Imogen:hip", "cfg must be a struct");
    imh_config c;
    std::memset(&c, 0, sizeof(c));
    double v[6];
    field_vec(s, "grid", v, 3);
    for (int i = 0; i < 3; ++i) c.grid[i] = (int)v[i];
    field_vec(s, "dgrid", c.dgrid, 3);
    c.gamma = field_or(s, "gamma", 5.0 / 3.0);
    c.minmass = field_or(s, "minmass", 1e-5);
    c.cfl = field_or(s, "cfl", 0.7);
    c.bc_infinity = (int)field_or(s, "bc_infinity", 20);
    field_vec(s, "bc", v, 6);  // 2 x 3, MATLAB column-major: v[side + 2*dim]
    for (int d = 0; d < 3; ++d)
        for (int side = 0; side < 2; ++side) c.bc[side][d] = (int)v[side + 2 * d];
    field_vec(s, "limiter", v, 3);
    for (int i = 0; i < 3; ++i) c.limiter[i] = (int)v[i];
    c.fluid_active = (int)field_or(s, "fluid_active", 1);
    c.magnet_active = (int)field_or(s, "magnet_active", 0);
    c.update_mode = (int)field_or(s, "update_mode", 1);
    c.iter_max = (int)field_or(s, "iter_max", 1000);
    c.time_max = field_or(s, "time_max", 5000);
    c.perms_order = (int)field_or(s, "perms_order", 0);
    c.precision = (int)field_or(s, "precision", IMH_FP64);
    c.device = (int)field_or(s, "device", 0);
    return c;
}

// The k-th component of a 3 x [grid] array or of a 1x3 cell array.
const double* component(const mxArray* a, int k, size_t n) {
    if (mxIsCell(a)) {
        const mxArray* c = mxGetCell(a, k);
        if (!c || !mxIsDouble(c) || mxGetNumberOfElements(c) != n)
            mexErrMsgIdAndTxt("Imogen:hip", "component %d has the wrong size or type", k + 1);
        return mxGetPr(c);
    }
    if (!mxIsDouble(a) || mxGetNumberOfElements(a) != 3 * n)
        mexErrMsgIdAndTxt("Imogen:hip", "expected 3 x grid doubles");
    // 3 x [grid] column-major interleaves components; copy them out
    static thread_local std::string buf[3];
    buf[k].resize(n * sizeof(double));
    double* dst = reinterpret_cast<double*>(&buf[k][0]);
    const double* src = mxGetPr(a);
    for (size_t i = 0; i < n; ++i) dst[i] = src[3 * i + k];
    return dst;
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    if (nrhs < 1 || !mxIsChar(prhs[0])) mexErrMsgIdAndTxt("Imogen:hip", "first argument: command");
    char* cmd_c = mxArrayToString(prhs[0]);
    std::string cmd(cmd_c);
    mxFree(cmd_c);
    if (cmd == "create") {
        if (nrhs < 2) mexErrMsgIdAndTxt("Imogen:hip", "create needs cfg");
        imh_config c = parse_config(prhs[1]);
        imh_ctx* h = nullptr;
        check(imh_create(&c, &h));
        plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
        *static_cast<uint64_t*>(mxGetData(plhs[0])) = (uint64_t)(uintptr_t)h;
        return;
    }
    if (nrhs < 2) mexErrMsgIdAndTxt("Imogen:hip", "%s needs a context handle", cmd.c_str());
    imh_ctx* h = handle(prhs[1]);
    if (cmd == "destroy") {
        check(imh_destroy(h));
    } else if (cmd == "upload") {
        if (nrhs < 6) mexErrMsgIdAndTxt("Imogen:hip", "upload(h, mass, mom, ener, mag)");
        size_t n = mxGetNumberOfElements(prhs[2]);
        const double* mom[3] = {component(prhs[3], 0, n), component(prhs[3], 1, n),
                                component(prhs[3], 2, n)};
        // copy mom before mag reuses the per-component buffers
        std::string keep[3];
        for (int k = 0; k < 3; ++k) {
            keep[k].assign(reinterpret_cast<const char*>(mom[k]), n * sizeof(double));
            mom[k] = reinterpret_cast<const double*>(keep[k].data());
        }
        const double* mag[3] = {component(prhs[5], 0, n), component(prhs[5], 1, n),
                                component(prhs[5], 2, n)};
        check(imh_upload_state(h, mxGetPr(prhs[2]), mom, mxGetPr(prhs[4]), mag, n));
    } else if (cmd == "fluid") {
        check(imh_fluid_sweep(h, (int)mxGetScalar(prhs[2])));
    } else if (cmd == "magnet") {
        const double* I = mxGetPr(prhs[3]);
        int I2[2] = {(int)I[0], (int)I[1]};
        check(imh_magnet_sweep(h, (int)mxGetScalar(prhs[2]), I2));
    } else if (cmd == "time_update") {
        check(imh_time_update(h, (int)mxGetScalar(prhs[2])));
    } else if (cmd == "flux") {
        check(imh_flux(h, (int)mxGetScalar(prhs[2])));
    } else if (cmd == "iteration") {
        check(imh_iteration(h));
    } else if (cmd == "run") {
        int done = 0;
        check(imh_run(h, (int)mxGetScalar(prhs[2]), &done));
        plhs[0] = mxCreateDoubleScalar(done);
    } else if (cmd == "source") {
        check(imh_source(h));
    } else if (cmd == "setGravity") {
        // setGravity(h, fixedPotential, MASS_THRESHOLD); an empty potential turns it off
        if (nrhs < 4) mexErrMsgIdAndTxt("Imogen:hip", "setGravity(h, potential, threshold)");
        size_t n = mxGetNumberOfElements(prhs[2]);
        check(imh_set_gravity(h, n ? mxGetPr(prhs[2]) : nullptr, n, mxGetScalar(prhs[3])));
    } else if (cmd == "setRadiation") {
        // setRadiation(h, type, exponent, initialMaximum); type 0 'empty', 1 'optically_thin'
        if (nrhs < 5) mexErrMsgIdAndTxt("Imogen:hip", "setRadiation(h, type, exponent, max)");
        check(imh_set_radiation(h, (int)mxGetScalar(prhs[2]), mxGetScalar(prhs[3]),
                                mxGetScalar(prhs[4])));
    } else if (cmd == "setViscosity") {
        // setViscosity(h, type, linear, quadratic); type 0 'empty', 1 'neumann_richtmyer'
        if (nrhs < 5) mexErrMsgIdAndTxt("Imogen:hip", "setViscosity(h, type, c1, c2)");
        check(imh_set_viscosity(h, (int)mxGetScalar(prhs[2]), mxGetScalar(prhs[3]),
                                mxGetScalar(prhs[4])));
    } else if (cmd == "download") {
        if (nrhs < 4) mexErrMsgIdAndTxt("Imogen:hip", "download(h, fieldId, numel)");
        size_t n = (size_t)mxGetScalar(prhs[3]);
        int field = (int)mxGetScalar(prhs[2]);
        plhs[0] = mxCreateDoubleMatrix((mwSize)n, 1, mxREAL);
        check(imh_download_field(h, field, mxGetPr(plhs[0]), n));
    } else if (cmd == "time") {
        double t, dt;
        int it, gi;
        check(imh_get_time(h, &t, &dt, &it, &gi));
        plhs[0] = mxCreateDoubleScalar(t);
        if (nlhs > 1) plhs[1] = mxCreateDoubleScalar(dt);
        if (nlhs > 2) plhs[2] = mxCreateDoubleScalar(it);
        if (nlhs > 3) plhs[3] = mxCreateDoubleScalar(gi);
    } else {
        mexErrMsgIdAndTxt("Imogen:hip", "unknown command '%s'", cmd.c_str());
    }
}
