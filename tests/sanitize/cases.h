// Geometries and case descriptors shared by the sanitizer drivers: the
// 64x32x32 channel, small synthetic TCPC junctions and random wall masks.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "lbm355.h"

struct Mask {
    int nx, ny, nz;
    std::vector<int32_t> v;  // x-fastest, 1 = wall
};

// straight channel along x, open within radius r of the y-z centre
inline Mask channel(int nx, int ny, int nz, double r)
{
    Mask m{nx, ny, nz, std::vector<int32_t>(size_t(nx) * ny * nz, 1)};
    const double cy = (ny - 1) / 2.0, cz = (nz - 1) / 2.0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                if ((y - cy) * (y - cy) + (z - cz) * (z - cz) <= r * r) m.v[x + size_t(nx) * (y + size_t(ny) * z)] = 0;
    return m;
}

// junction: IVC along x from x-min, SVC from x-max (axes `off` apart in y),
// pulmonary tube along y through the centre
inline Mask junction(int X, int Y, int Z, double ri, double rs, double rp, double off)
{
    Mask m{X, Y, Z, std::vector<int32_t>(size_t(X) * Y * Z, 1)};
    const int xc = X / 2, zc = Z / 2;
    const double yi = Y / 2 - off / 2.0, ys = Y / 2 + off / 2.0;
    for (int z = 0; z < Z; ++z)
        for (int y = 0; y < Y; ++y)
            for (int x = 0; x < X; ++x) {
                const bool ivc = (y - yi) * (y - yi) + double(z - zc) * (z - zc) <= ri * ri && x <= xc;
                const bool svc = (y - ys) * (y - ys) + double(z - zc) * (z - zc) <= rs * rs && x >= xc;
                const bool pa = double(x - xc) * (x - xc) + double(z - zc) * (z - zc) <= rp * rp;
                if (ivc || svc || pa) m.v[x + size_t(X) * (y + size_t(Y) * z)] = 0;
            }
    return m;
}

// random wall mask with open fraction p (the z planes are walls anyway)
inline Mask random_mask(std::mt19937& rng, int nx, int ny, int nz, double p)
{
    Mask m{nx, ny, nz, std::vector<int32_t>(size_t(nx) * ny * nz)};
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (auto& c : m.v) c = u(rng) < p ? 0 : 1;
    return m;
}

// every cell open
inline Mask random_mask_free(int nx, int ny, int nz)
{
    return Mask{nx, ny, nz, std::vector<int32_t>(size_t(nx) * ny * nz, 0)};
}

inline lbm355_case_desc describe(const Mask& m, int collision, int storage)
{
    lbm355_case_desc d;
    std::memset(&d, 0, sizeof d);
    d.nx = m.nx; d.ny = m.ny; d.nz = m.nz;
    d.geometry = m.v.data();
    d.dl = 1e-3;
    d.dt = 1e-3;
    d.collision = collision;
    d.storage = storage;
    d.nu_ref = 0.02;
    d.profile = LBM355_PROFILE_PLUG;
    d.ivc_flux = 2.0;
    d.svc_flux = 1.5;
    d.ivc_center[0] = d.ivc_center[1] = -1.0;
    d.svc_center[0] = d.svc_center[1] = -1.0;
    d.stats_every = 1;
    return d;
}

inline void carreau_yasuda(lbm355_case_desc& d, int model)
{
    d.cy_enabled = 1;
    d.cy_model = model;
    d.cy_nu0 = 0.3;
    d.cy_lambda = 50.0;
    d.cy_a = 2.0;
    d.cy_n = 0.3568;
}

inline void channel_faces(lbm355_case_desc& d)
{
    const int f[6] = {LBM355_FACE_INFLOW, LBM355_FACE_OUTFLOW, LBM355_FACE_WALL, LBM355_FACE_WALL, LBM355_FACE_WALL,
                      LBM355_FACE_WALL};
    std::memcpy(d.face_bc, f, sizeof f);
}
