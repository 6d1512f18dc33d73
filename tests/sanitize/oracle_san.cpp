// Host sanitizer run of the CPU oracle (oracle/lbm_oracle.cpp), built with
// -fsanitize=address,undefined: every collision model in both storage
// types, the three Carreau-Yasuda couplings, statistics, strain and every
// field read, population round trips and the streaming-only path, on the
// channel and on junctions whose brick origins shift.  Exit status 0 and
// "oracle_san: ok" when no check or sanitizer fired.
#include <cstdio>
#include <vector>

#include "cases.h"

extern "C" {
void* oracle_create(const lbm355_case_desc* d);
void oracle_destroy(void* o);
int oracle_step(void* o, int64_t n);
int oracle_set_inflow(void* o, double a, double b);
int oracle_get_field(void* o, int field, double* out);
int oracle_get_populations(void* o, double* out);
int oracle_set_populations(void* o, const double* in);
int oracle_compute_strain(void* o);
int oracle_stream_only(void* o, int64_t n);
int oracle_info(void* o, lbm355_case_info* i);
const char* oracle_last_error(void);
}

static int failures = 0;

static void check(bool ok, const char* what)
{
    if (!ok) {
        std::fprintf(stderr, "oracle_san: %s failed (%s)\n", what, oracle_last_error());
        ++failures;
    }
}

static void run(const Mask& m, lbm355_case_desc d, int steps)
{
    void* o = oracle_create(&d);
    check(o != nullptr, "create");
    if (!o) return;
    const size_t n = size_t(m.nx) * m.ny * m.nz;
    check(oracle_step(o, steps) == 0, "step");
    check(oracle_set_inflow(o, 0.7 * d.ivc_flux, 0.9 * d.svc_flux) == 0, "set_inflow");
    check(oracle_step(o, 3) == 0, "step after set_inflow");
    check(oracle_compute_strain(o) == 0, "compute_strain");
    std::vector<double> f(n);
    for (int k = 0; k < LBM355_FIELD_COUNT; ++k) {
        check(oracle_get_field(o, k, f.data()) == 0, "get_field");
        for (double v : f)
            if (!std::isfinite(v)) { check(false, "finite field"); break; }
    }
    std::vector<double> pops(27 * n);
    check(oracle_get_populations(o, pops.data()) == 0, "get_populations");
    check(oracle_set_populations(o, pops.data()) == 0, "set_populations");
    bool open_faces = false;
    for (int k = 0; k < 6; ++k) open_faces = open_faces || d.face_bc[k] != LBM355_FACE_WALL;
    if (!open_faces) check(oracle_stream_only(o, 2) == 0, "stream_only");
    lbm355_case_info info;
    check(oracle_info(o, &info) == 0, "info");
    oracle_destroy(o);
}

int main()
{
    const Mask ch = channel(64, 32, 32, 12.0);
    const Mask jn = junction(40, 30, 20, 5.0, 4.0, 3.5, 2.0);
    const Mask odd = junction(37, 29, 11, 4.0, 3.0, 3.0, 1.0);   // sizes that are no multiple of the brick
    const int models[4] = {LBM355_CUM, LBM355_BGK, LBM355_MRT, LBM355_KBC};
    for (int st : {LBM355_F64, LBM355_F32})
        for (int c : models) {
            lbm355_case_desc d = describe(ch, c, st);
            channel_faces(d);
            d.profile = LBM355_PROFILE_PARABOLIC;
            d.ivc_umax = 0.05;
            d.ivc_radius = 12.0;
            d.ivc_center[0] = d.ivc_center[1] = 15.5;
            run(ch, d, 4);
            run(jn, describe(jn, c, st), 4);
        }
    for (int cy : {LBM355_CY_OFF, LBM355_CY_NEQ, LBM355_CY_FD_FORCE}) {
        lbm355_case_desc d = describe(odd, LBM355_CUM, LBM355_F64);
        carreau_yasuda(d, cy);
        d.stats_every = 3;
        d.pulse_amp = 0.3;
        d.pulse_period = 40.0;
        run(odd, d, 5);
    }
    // closed box: every face a wall (the streaming-only path)
    {
        const Mask box = random_mask_free(24, 17, 9);
        lbm355_case_desc d = describe(box, LBM355_BGK, LBM355_F64);
        for (int k = 0; k < 6; ++k) d.face_bc[k] = LBM355_FACE_WALL;
        run(box, d, 3);
    }
    if (failures) return 1;
    std::printf("oracle_san: ok\n");
    return 0;
}
