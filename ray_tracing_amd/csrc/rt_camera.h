// rt_camera.h -- device camera models for ray generation and reprojection.
//
// pinhole32: Newton undistortion in FP64, programs.cu:308-373
// equi62_lut1d: one-step tangential removal + 1024-entry theta LUT in FP64,
//               programs.cu:375-515
// transform_direction: programs.cu:521-532
#pragma once

#include "rt_internal.h"
#include "rt_math.h"

namespace rtk {

struct d2 {
    double x, y;
};

__device__ __forceinline__ d2 distort_pinhole(const rt_camera& c, d2 xy) {
    const double x = xy.x, y = xy.y;
    const double x2 = x * x, y2 = y * y, xyt = x * y;
    const double r2 = x2 + y2, r4 = r2 * r2, r6 = r4 * r2;
    const double radial = 1.0 + c.k1 * r2 + c.k2 * r4 + c.k3 * r6;
    return d2{x * radial + c.p2 * (r2 + 2.0 * x2) + 2.0 * c.p1 * xyt,
        y * radial + c.p1 * (r2 + 2.0 * y2) + 2.0 * c.p2 * xyt};
}

__device__ __forceinline__ bool solve2(double a00, double a01, double a10, double a11, d2 rhs,
    d2* delta) {
    const double det = a00 * a11 - a01 * a10;
    if (fabs(det) < 1e-12) return false;
    const double inv = 1.0 / det;
    delta->x = (a11 * rhs.x - a01 * rhs.y) * inv;
    delta->y = (-a10 * rhs.x + a00 * rhs.y) * inv;
    return true;
}

__device__ inline f3 unproject_pinhole(const rt_camera& c, double px, double py) {
    const d2 xyd{(px - c.cx) / c.fx, (py - c.cy) / c.fy};
    d2 xy = xyd;
    for (int it = 0; it < 24; ++it) {
        const d2 dist = distort_pinhole(c, xy);
        const d2 err{xyd.x - dist.x, xyd.y - dist.y};
        if (err.x * err.x + err.y * err.y < 1e-20) break;
        const double x = xy.x, y = xy.y;
        const double r2 = x * x + y * y, r4 = r2 * r2, r6 = r4 * r2;
        const double radial = 1.0 + c.k1 * r2 + c.k2 * r4 + c.k3 * r6;
        const double drdx = 2.0 * x * (c.k1 + 2.0 * c.k2 * r2 + 3.0 * c.k3 * r4);
        const double drdy = 2.0 * y * (c.k1 + 2.0 * c.k2 * r2 + 3.0 * c.k3 * r4);
        d2 delta;
        if (!solve2(radial + x * drdx + 6.0 * c.p2 * x + 2.0 * c.p1 * y,
                x * drdy + 2.0 * c.p1 * x + 2.0 * c.p2 * y,
                y * drdx + 2.0 * c.p1 * x + 2.0 * c.p2 * y,
                radial + y * drdy + 6.0 * c.p1 * y + 2.0 * c.p2 * x, err, &delta))
            break;
        xy.x += delta.x;
        xy.y += delta.y;
    }
    return normalize3(mk3((float)xy.x, (float)xy.y, 1.0f));
}

__device__ __forceinline__ d2 equi_tangential(d2 xy, const double* tg) {
    const double xr = xy.x, yr = xy.y, xr2 = xr * xr, yr2 = yr * yr, xryr = xr * yr;
    return d2{xr + 2.0 * tg[0] * xryr + tg[1] * (xr2 + yr2 + 2.0 * xr2),
        yr + tg[0] * (xr2 + yr2 + 2.0 * yr2) + 2.0 * tg[1] * xryr};
}

__device__ inline f3 unproject_equi(const rt_camera& c, double px, double py) {
    const d2 xy{(px - c.cx) / c.fx, (py - c.cy) / c.fy};
    if (xy.x * xy.x + xy.y * xy.y < 1e-12 * 1e-12) return mk3(0.0f, 0.0f, 1.0f);
    const double* tg = c.tangential;
    const d2 dist = equi_tangential(xy, tg);
    d2 delta;
    d2 xr = xy;
    if (solve2(1.0 + 2.0 * tg[0] * xy.y + 6.0 * tg[1] * xy.x,
            2.0 * tg[0] * xy.x + 2.0 * tg[1] * xy.y, 2.0 * tg[0] * xy.x + 2.0 * tg[1] * xy.y,
            1.0 + 6.0 * tg[0] * xy.y + 2.0 * tg[1] * xy.x, d2{xy.x - dist.x, xy.y - dist.y},
            &delta)) {
        xr = d2{xy.x + delta.x, xy.y + delta.y};
    }
    const double rd = sqrt(xr.x * xr.x + xr.y * xr.y);
    if (rd < 1e-12) return mk3(0.0f, 0.0f, 1.0f);
    double theta = 0.0;
    if (!(rd <= 0.0 || c.lut_step <= 0.0)) {
        const double pos = rd / c.lut_step;
        if (pos > 1023.0) return normalize3(mk3((float)xy.x, (float)xy.y, 1.0f));
        if (pos >= 1023.0) {
            theta = c.lut[1023];
        } else {
            const int idx = (int)pos;
            const double alpha = pos - (double)idx;
            theta = (1.0 - alpha) * c.lut[idx] + alpha * c.lut[idx + 1];
        }
    }
    const double scale = tan(theta) / rd;
    return normalize3(mk3((float)(xr.x * scale), (float)(xr.y * scale), 1.0f));
}

// project_pinhole32 / project_equi62_lut1d (programs.cu:375-473), camera-space
// direction -> pixel
__device__ inline void project_camera(const rt_camera& c, f3 dc, float* px, float* py) {
    if (c.model == RT_CAM_EQUI62_LUT1D) {
        const double dx = dc.x, dy = dc.y, dz = dc.z;
        const double r = sqrt(dx * dx + dy * dy);
        if (r < 1e-12) {
            *px = (float)c.cx;
            *py = (float)c.cy;
            return;
        }
        const double theta = atan2(r, dz);
        double rd;
        if (theta <= 0.0 || c.lut_step <= 0.0) {
            rd = 0.0;
        } else if (theta >= c.lut[1023]) {
            rd = 1023.0 * c.lut_step;
        } else {
            // first LUT interval containing theta (the LUT is non-decreasing)
            int lo = 0, hi = 1022;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (c.lut[mid + 1] < theta) lo = mid + 1;
                else hi = mid;
            }
            while (lo > 0 && c.lut[lo] > theta) --lo;
            rd = 1023.0 * c.lut_step;
            for (int i = lo; i < 1023; ++i) {
                if (c.lut[i] <= theta && theta <= c.lut[i + 1]) {
                    const double den = c.lut[i + 1] - c.lut[i];
                    const double a = den > 0.0 ? (theta - c.lut[i]) / den : 0.0;
                    rd = ((double)i + a) * c.lut_step;
                    break;
                }
            }
        }
        const d2 xyr{dx / r * rd, dy / r * rd};
        const d2 xyd = equi_tangential(xyr, c.tangential);
        *px = (float)(xyd.x * c.fx + c.cx);
        *py = (float)(xyd.y * c.fy + c.cy);
        return;
    }
    if (fabsf(dc.z) < 1e-7f) {
        *px = -1.0f;
        *py = -1.0f;
        return;
    }
    const double inv_z = 1.0 / (double)dc.z;
    const d2 xyd = distort_pinhole(c, d2{(double)dc.x * inv_z, (double)dc.y * inv_z});
    *px = (float)(xyd.x * c.fx + c.cx);
    *py = (float)(xyd.y * c.fy + c.cy);
}

__device__ __forceinline__ f3 camera_transform(const rt_camera& c, f3 dc) {
    const f3 bx = mk3((float)c.basis_x[0], (float)c.basis_x[1], (float)c.basis_x[2]);
    const f3 by = mk3((float)c.basis_y[0], (float)c.basis_y[1], (float)c.basis_y[2]);
    const f3 bz = mk3((float)c.basis_z[0], (float)c.basis_z[1], (float)c.basis_z[2]);
    return normalize3((bx * dc.x + by * dc.y) + bz * dc.z);
}

__device__ __forceinline__ void camera_ray(const rt_camera& c, double px, double py, f3* o, f3* d) {
    const f3 dc = c.model == RT_CAM_EQUI62_LUT1D ? unproject_equi(c, px, py)
                                                 : unproject_pinhole(c, px, py);
    *d = camera_transform(c, dc);
    *o = mk3((float)c.origin[0], (float)c.origin[1], (float)c.origin[2]);
}

}  // namespace rtk
