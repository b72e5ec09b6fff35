// bonded.h — bonded force routines on the device, FP32 with the reference's
// FP64 intermediates and exact operation order (FMA contraction disabled).
//   calcPairbondForces     Engine.cu:525-548
//   calcAnglebondForces    Engine.cu:552-578
//   calcDihedralbondForces Engine.cu:580-637
//   Float3 norm/len/cross/dot/getAngle  QuantomTypes.cuh:50-90
#pragma once
#include "device_common.h"

namespace lima {

struct F3 {
    float x, y, z;
};

LIMA_DEV F3 f3(float x, float y, float z) { return F3{x, y, z}; }
LIMA_DEV F3 f3_add(F3 a, F3 b) { return F3{fadd(a.x, b.x), fadd(a.y, b.y), fadd(a.z, b.z)}; }
LIMA_DEV F3 f3_sub(F3 a, F3 b) { return F3{fsub(a.x, b.x), fsub(a.y, b.y), fsub(a.z, b.z)}; }
// Float3::operator*(double): FP64 product narrowed to FP32 (QuantomTypes.cuh:39)
LIMA_DEV F3 f3_muld(F3 a, double s) { return F3{(float)((double)a.x * s), (float)((double)a.y * s), (float)((double)a.z * s)}; }
LIMA_DEV float f3_len(F3 a) { return exact_len(a.x, a.y, a.z); }
LIMA_DEV F3 f3_norm(F3 a) {
    float l = f3_len(a);
    if (l != 0.0f) return f3_muld(a, (double)__fdiv_rn(1.0f, l));
    return F3{0.0f, 0.0f, 0.0f};
}
LIMA_DEV F3 f3_cross(F3 a, F3 b) {
    return F3{fsub(fmul(a.y, b.z), fmul(a.z, b.y)), fsub(fmul(a.z, b.x), fmul(a.x, b.z)),
              fsub(fmul(a.x, b.y), fmul(a.y, b.x))};
}
LIMA_DEV float f3_dot(F3 a, F3 b) { return fadd(fadd(fmul(a.x, b.x), fmul(a.y, b.y)), fmul(a.z, b.z)); }
LIMA_DEV F3 f3_neg(F3 a) { return F3{-a.x, -a.y, -a.z}; }

LIMA_DEV float get_angle(F3 v1, F3 v2) {
    float val = __fdiv_rn(f3_dot(v1, v2), fmul(f3_len(v1), f3_len(v2)));
    val = val > 1.0f ? 1.0f : val;
    val = val < -1.0f ? -1.0f : val;
    return acosf(val);
}

LIMA_DEV void pairbond(F3 a, F3 b, float b0, float kb, F3* res, double* potE) {
    F3 difference = f3_sub(a, b);
    double error = (double)fsub(f3_len(difference), b0);
    *potE += 0.5 * (double)kb * (error * error);
    difference = f3_norm(difference);
    double force_scalar = -(double)kb * error;
    res[0] = f3_muld(difference, force_scalar);
    res[1] = f3_muld(f3_muld(difference, force_scalar), -1.0);
}

LIMA_DEV void anglebond(F3 l, F3 m, F3 r, float theta0, float k, F3* res, double* potE) {
    F3 v1 = f3_sub(l, m);
    F3 v2 = f3_sub(r, m);
    F3 normal1 = f3_cross(v1, v2);
    F3 normal2 = f3_cross(v2, v1);
    F3 dir1 = f3_norm(f3_cross(v1, normal2));
    F3 dir2 = f3_norm(f3_cross(v2, normal1));
    double angle = (double)get_angle(v1, v2);
    double error = angle - (double)theta0;
    *potE += 0.5 * (double)k * error * error * 0.5;
    double force_scalar = (double)k * error;
    res[0] = f3_muld(dir1, force_scalar);
    res[2] = f3_muld(dir2, force_scalar);
    res[1] = f3_muld(f3_add(res[0], res[2]), -1.0);
}

LIMA_DEV void dihedral(F3 l, F3 lm, F3 rm, F3 r, float phi0, float k, int n, F3* res, double* potE) {
    F3 normal1 = f3_norm(f3_cross(f3_sub(l, lm), f3_sub(rm, lm)));
    F3 normal2 = f3_norm(f3_cross(f3_sub(lm, rm), f3_sub(r, rm)));
    float torsion = get_angle(normal1, normal2);
    if (f3_len(f3_sub(f3_add(l, normal1), r)) < f3_len(f3_sub(f3_sub(l, normal1), r)))
        normal2 = f3_neg(normal2);
    else
        normal1 = f3_neg(normal1);
    float arg = fsub(fmul((float)n, torsion), phi0);
    float force_scalar = fmul(k, sinf(arg));
    *potE += (double)fmul(k, fsub(1.0f, cosf(arg)));
    res[0] = f3_muld(normal1, (double)force_scalar);
    res[3] = f3_muld(normal2, (double)force_scalar);
    res[1] = f3_muld(f3_muld(f3_add(res[0], res[3]), -1.0), 0.5);
    res[2] = res[1];
}

// device term records (compound- or bridge-local indices)
struct BondD {
    float b0, kb;
    int a0, a1;
};
struct AngleD {
    float t0, k;
    int a0, a1, a2, pad;
};
struct DihD {
    float p0, k;
    int n, a0, a1, a2, a3, pad;
};

}  // namespace lima
