// mpm_device.h -- device math of the explicit step: uGIMP 1D factors, 3x3 algebra and the
// constitutive laws, written for gfx950 (FP64 VALU). Every formula keeps the reference's order
// of floating-point operations so results agree to rounding; the code is built with
// -ffp-contract=off so no multiply-add is fused where the reference's x86 build does not fuse.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/nairnmpm_gpu.h"

#define MPM_SQRT_TWOTHIRDS 0.8164965809277260   // System/MPMPrefix.hpp:92

// ---------------------------------------------------------------------------------------
// uGIMP, factorised per axis. A particle in a cell touches nodes at offsets -1..2 from the
// cell's lower node on each axis; the reference enumerates the 16 (2D) or 64 (3D) candidates
// (Common/Elements/FourNodeIsoparam.cpp:27-31, Elements/EightNodeIsoparamBrick.cpp:23-53) and
// keeps those with |xi - g| < 2 + lp on every axis. The per-axis factor for offset o is a
// function of xp = |xi - (2o-1)| only, so it is computed once per axis and offset here.
// S < 0 marks a candidate outside the support on that axis.

// 3D per-axis factor (EightNodeIsoparamBrick.cpp:327-389). inv_size is 1/(4 lp) of the axis,
// except z, where the reference uses lp.y (line 303).
__device__ __forceinline__ void gimp_axis3(double xi, double lp, double inv_size, double S[4],
                                           double dS[4]) {
    double q1 = 2. - lp, q2 = 2. + lp;
#pragma unroll
    for (int o = 0; o < 4; o++) {
        double g = 2. * (double)(o - 1) - 1.;
        double xp = fabs(xi - g);
        if (xp >= q2) { S[o] = -1.; dS[o] = 0.; continue; }
        double s, d, arg;
        if (xp < lp) { s = ((4. - lp) * lp - xp * xp) * inv_size; d = -xp / (2. * lp); }
        else if (xp <= q1) { s = 0.5 * (2. - xp); d = -0.5; }
        else { arg = (q2 - xp) * inv_size; s = 2. * lp * arg * arg; d = -arg; }
        double sign = xi > g ? 1. : -1.;
        S[o] = s;
        dS[o] = sign * d;
    }
}

// 2D per-axis factor (FourNodeIsoparam.cpp:462-510)
__device__ __forceinline__ void gimp_axis2(double xi, double lp, double S[4], double dS[4]) {
    double q1 = 2. - lp, q2 = 2. + lp;
    double inv_size = 1. / (4. * lp);
#pragma unroll
    for (int o = 0; o < 4; o++) {
        double g = 2. * (double)(o - 1) - 1.;
        double xp = fabs(xi - g);
        if (xp >= q2) { S[o] = -1.; dS[o] = 0.; continue; }
        double s, d, arg;
        if (xp < lp) { s = ((4. - lp) * lp - xp * xp) * inv_size; d = -xp * inv_size * 2.0; }
        else if (xp <= q1) { s = (2. - xp) / 2.; d = -0.5; }
        else { arg = (q2 - xp) * inv_size; s = 2. * lp * arg * arg; d = -arg; }
        double sign = xi > g ? 1. : -1.;
        S[o] = s;
        dS[o] = sign * d;
    }
}

// classic MPM (ElementBase::useGimp = POINT_GIMP): the element's linear functions on one axis,
// (1 - xi)/2 and (1 + xi)/2 at its two nodes, so that their products are the reference's
// 0.25*(1+xii*xi)*(1+eti*eta) and 0.125*(...)(...)(...) bit for bit (FourNodeIsoparam.cpp:189-214,
// EightNodeIsoparamBrick.cpp:87-108); candidates -1 and 2 are not nodes of the particle.
__device__ __forceinline__ void linear_axis(double xi, double S[4], double dS[4]) {
    S[0] = -1.; dS[0] = 0.;
    S[1] = 0.5 * (1. - xi); dS[1] = -0.5;
    S[2] = 0.5 * (1. + xi); dS[2] = 0.5;
    S[3] = -1.; dS[3] = 0.;
}


// ---------------------------------------------------------------------------------------
// 3x3 matrix with the reference's 2D flag semantics (Common/System/Matrix3.cpp)
struct DMat3 {
    double m[3][3];
};

// F from the particle's strain/rotation encoding (MatPoint3D.cpp:364-377)
__device__ __forceinline__ void get_F3(const double ep[6], const double wr[3], double F[3][3]) {
    // ep: xx,yy,zz,yz,xz,xy ; wr: yz,xz,xy
    F[0][0] = 1. + ep[0];
    F[1][1] = 1. + ep[1];
    F[2][2] = 1. + ep[2];
    F[0][1] = 0.5 * (ep[5] - wr[2]);
    F[1][0] = 0.5 * (ep[5] + wr[2]);
    F[0][2] = 0.5 * (ep[4] - wr[1]);
    F[2][0] = 0.5 * (ep[4] + wr[1]);
    F[1][2] = 0.5 * (ep[3] - wr[0]);
    F[2][1] = 0.5 * (ep[3] + wr[0]);
}
// MatPoint3D::SetDeformationGradientMatrix (MatPoint3D.cpp:320-336)
__device__ __forceinline__ void set_F3(const double F[3][3], double ep[6], double wr[3]) {
    ep[0] = F[0][0] - 1.;
    ep[1] = F[1][1] - 1.;
    ep[2] = F[2][2] - 1.;
    ep[5] = F[1][0] + F[0][1];
    ep[4] = F[2][0] + F[0][2];
    ep[3] = F[2][1] + F[1][2];
    wr[2] = F[1][0] - F[0][1];
    wr[1] = F[2][0] - F[0][2];
    wr[0] = F[2][1] - F[1][2];
}
// general 3x3 product, reference summation order (Matrix3.cpp:948-962)
__device__ __forceinline__ void mat_mul3(const double a[3][3], const double b[3][3], double c[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double s = 0.;
#pragma unroll
            for (int k = 0; k < 3; k++) s += a[i][k] * b[k][j];
            c[i][j] = s;
        }
}
// 2D product (Matrix3.cpp:924-944); entries [0..1][0..1] and [2][2]
__device__ __forceinline__ void mat_mul2(const double a[3][3], const double b[3][3], double c[3][3]) {
    double m00 = a[0][0] * b[0][0] + a[0][1] * b[1][0];
    double m01 = a[0][0] * b[0][1] + a[0][1] * b[1][1];
    double m10 = a[1][0] * b[0][0] + a[1][1] * b[1][0];
    double m11 = a[1][0] * b[0][1] + a[1][1] * b[1][1];
    c[0][0] = m00; c[0][1] = m01; c[0][2] = 0.;
    c[1][0] = m10; c[1][1] = m11; c[1][2] = 0.;
    c[2][0] = 0.; c[2][1] = 0.; c[2][2] = a[2][2] * b[2][2];
}

// Particle state of the constitutive update, in registers
struct PState {
    double ep[6], wr[3], sp[6], P, alt[6], h0, h1;
};

// ---------------------------------------------------------------------------------------
// Elastic::HypoIncrementDeformation (ElasticMPM.cpp:392-403) with Exponential(1)
template <int DIM>
__device__ __forceinline__ void hypo_increment(const double du[3][3], PState &s) {
    double dF[3][3], pF[3][3], F[3][3];
    if (DIM == 3) {
        dF[0][0] = 1. + du[0][0]; dF[0][1] = du[0][1]; dF[0][2] = du[0][2];
        dF[1][0] = du[1][0]; dF[1][1] = 1. + du[1][1]; dF[1][2] = du[1][2];
        dF[2][0] = du[2][0]; dF[2][1] = du[2][1]; dF[2][2] = 1. + du[2][2];
        get_F3(s.ep, s.wr, pF);
        mat_mul3(dF, pF, F);
        set_F3(F, s.ep, s.wr);
    } else {
        dF[0][0] = 1. + du[0][0]; dF[0][1] = du[0][1]; dF[1][0] = du[1][0]; dF[1][1] = 1. + du[1][1];
        dF[2][2] = 1. + du[2][2]; dF[0][2] = dF[1][2] = dF[2][0] = dF[2][1] = 0.;
        // MatPoint2D::GetDeformationGradient (MatPoint2D.cpp:369-377)
        pF[0][0] = 1. + s.ep[0]; pF[0][1] = 0.5 * (s.ep[5] - s.wr[2]);
        pF[1][0] = 0.5 * (s.ep[5] + s.wr[2]); pF[1][1] = 1. + s.ep[1]; pF[2][2] = 1. + s.ep[2];
        pF[0][2] = pF[1][2] = pF[2][0] = pF[2][1] = 0.;
        mat_mul2(dF, pF, F);
        // MatPoint2D::SetDeformationGradientMatrix (MatPoint2D.cpp:310-322)
        s.ep[0] = F[0][0] - 1.; s.ep[1] = F[1][1] - 1.; s.ep[2] = F[2][2] - 1.;
        s.ep[5] = F[1][0] + F[0][1];
        s.wr[2] = F[1][0] - F[0][1];
    }
}

// IsotropicMat small-rotation law (MoreIsotropicMat.cpp:185-351) + Hypo2D/3DCalculations
// (MaterialBaseMPM.cpp:1012-1047), no residual strains
template <int DIM>
__device__ __forceinline__ void law_isotropic(const double du[3][3], const MpmMaterial &m, PState &s) {
    hypo_increment<DIM>(du, s);
    if (DIM == 3) {
        double dvxx = du[0][0], dvyy = du[1][1], dvzz = du[2][2];
        double dgamxy = du[0][1] + du[1][0];
        double dgamxz = du[0][2] + du[2][0];
        double dgamyz = du[1][2] + du[2][1];
        double dwxy = du[1][0] - du[0][1];
        double dwxz = du[2][0] - du[0][2];
        double dwyz = du[2][1] - du[1][2];
        double d0 = m.C[0][0] * dvxx + m.C[0][1] * dvyy + m.C[0][2] * dvzz;
        double d1 = m.C[1][0] * dvxx + m.C[1][1] * dvyy + m.C[1][2] * dvzz;
        double d2 = m.C[2][0] * dvxx + m.C[2][1] * dvyy + m.C[2][2] * dvzz;
        double d3 = m.C[3][3] * dgamyz;
        double d4 = m.C[4][4] * dgamxz;
        double d5 = m.C[5][5] * dgamxy;
        const double *sp = s.sp;
        double st0 = -dwxy * sp[5] - dwxz * sp[4];
        double st1 = dwxy * sp[5] - dwyz * sp[3];
        double st2 = dwxz * sp[4] + dwyz * sp[3];
        double st3 = 0.5 * (dwxy * sp[4] + dwxz * sp[5] + dwyz * (sp[1] - sp[2]));
        double st4 = 0.5 * (-dwxy * sp[3] + dwxz * (sp[0] - sp[2]) + dwyz * sp[5]);
        double st5 = 0.5 * (dwxy * (sp[0] - sp[1]) - dwxz * sp[3] - dwyz * sp[4]);
        s.sp[0] += d0 + st0; s.sp[1] += d1 + st1; s.sp[2] += d2 + st2;
        s.sp[3] += d3 + st3; s.sp[4] += d4 + st4; s.sp[5] += d5 + st5;
    } else {
        double dvxx = du[0][0], dvyy = du[1][1];
        double dgam = du[0][1] + du[1][0];
        double dwrotxy = du[1][0] - du[0][1];
        double c1 = m.C[1][1] * dvxx + m.C[1][2] * dvyy;
        double c2 = m.C[1][2] * dvxx + m.C[2][2] * dvyy;
        double c3 = m.C[3][3] * dgam;
        double dnorm = dwrotxy * s.sp[5];
        double dshear = 0.5 * dwrotxy * (s.sp[0] - s.sp[1]);
        s.sp[0] += c1 - dnorm;
        s.sp[1] += c2 + dnorm;
        s.sp[5] += c3 + dshear;
        s.sp[2] += m.C[4][1] * dvxx + m.C[4][2] * dvyy + m.C[4][4] * 0.;
    }
}

// IsoPlasticity + LinearHardening (IsoPlasticity.cpp:128-451, LinearHardening.cpp:93-147)
template <int DIM>
__device__ __forceinline__ void law_isoplastic(const double de[3][3], const MpmMaterial &m, PState &s) {
    hypo_increment<DIM>(de, s);
    const bool is2D = DIM == 2;
    double delV = (de[0][0] + de[1][1] + de[2][2]) - 3. * 0.;
    double dgxy = de[0][1] + de[1][0], dgxz = 0., dgyz = 0.;
    if (!is2D) { dgxz = de[0][2] + de[2][0]; dgyz = de[1][2] + de[2][1]; }
    double dexxr = de[0][0], deyyr = de[1][1], dezzr = de[2][2];
    double dP = -m.Kred * delV;
    s.P += dP;
    double exx = s.alt[0], eyy = s.alt[1], ezz = s.alt[2], eyz = s.alt[3], exz = s.alt[4], exy = s.alt[5];
    double sxx = s.sp[0], syy = s.sp[1], szz = s.sp[2], syz = s.sp[3], sxz = s.sp[4], sxy = s.sp[5];
    double dwrotxy = de[1][0] - de[0][1];
    double st0[6] = {sxx, syy, szz, syz, sxz, sxy};
    if (is2D) {
        double dnorm = 0.5 * dwrotxy * exy;
        s.alt[0] -= dnorm;
        s.alt[1] += dnorm;
        s.alt[5] += dwrotxy * (exx - eyy);
        double dn = dwrotxy * sxy;
        st0[0] -= dn;
        st0[1] += dn;
        st0[5] += 0.5 * dwrotxy * (exx - eyy);     // IsoPlasticity.cpp:252 uses ep, not sp
    } else {
        double dwrotxz = de[2][0] - de[0][2];
        double dwrotyz = de[2][1] - de[1][2];
        double dxy = 0.5 * dwrotxy * exy;
        double dxz = 0.5 * dwrotxz * exz;
        double dyz = 0.5 * dwrotyz * eyz;
        s.alt[0] += -dxy - dxz;
        s.alt[1] += dxy - dyz;
        s.alt[2] += dxz + dyz;
        s.alt[3] += dwrotyz * (eyy - ezz) + 0.5 * (dwrotxz * exy + dwrotxy * exz);
        s.alt[4] += dwrotxz * (exx - ezz) + 0.5 * (dwrotyz * exy - dwrotxy * eyz);
        s.alt[5] += dwrotxy * (exx - eyy) - 0.5 * (dwrotyz * exz + dwrotxz * eyz);
        double sdxy = dwrotxy * sxy;
        double sdxz = dwrotxz * sxz;
        double sdyz = dwrotyz * syz;
        st0[0] += -sdxy - sdxz;
        st0[1] += sdxy - sdyz;
        st0[2] += sdxz + sdyz;
        st0[3] += 0.5 * (dwrotyz * (syy - szz) + dwrotxz * sxy + dwrotxy * sxz);
        st0[4] += 0.5 * (dwrotxz * (sxx - szz) + dwrotyz * sxy - dwrotxy * syz);
        st0[5] += 0.5 * (dwrotxy * (sxx - syy) - dwrotyz * sxz - dwrotxz * syz);
    }
    double thirdDelV = delV / 3.;
    double dels[6] = {0., 0., 0., 0., 0., 0.};
    dels[0] = 2. * m.Gred * (dexxr - thirdDelV);
    dels[1] = 2. * m.Gred * (deyyr - thirdDelV);
    dels[5] = m.Gred * dgxy;
    dels[2] = 2. * m.Gred * (dezzr - thirdDelV);
    if (!is2D) { dels[3] = m.Gred * dgyz; dels[4] = m.Gred * dgxz; }
    double str[6];
#pragma unroll
    for (int c = 0; c < 6; c++) str[c] = st0[c] + dels[c];
    if (is2D) { str[3] = 0.; str[4] = 0.; }
    double alpint = s.h0;
    double sq = str[0] * str[0] + str[1] * str[1] + str[2] * str[2];
    double t = is2D ? str[5] * str[5] : str[5] * str[5] + str[4] * str[4] + str[3] * str[3];
    double strialmag = sqrt(sq + t + t);
    double yld = alpint < m.alphaMax ? m.yldred + m.Epred * alpint : m.yldredMin;
    double ftrial = strialmag - MPM_SQRT_TWOTHIRDS * yld;
    if (ftrial < 0.) {
#pragma unroll
        for (int c = 0; c < 6; c++) s.sp[c] = str[c];
        return;
    }
    double lambdak = (strialmag - MPM_SQRT_TWOTHIRDS * (m.yldred + m.Epred * alpint)) / (2. * (m.Gred + m.Epred / 3.));
    if (alpint + MPM_SQRT_TWOTHIRDS * lambdak > m.alphaMax)
        lambdak = (strialmag - MPM_SQRT_TWOTHIRDS * m.yldredMin) / (2. * m.Gred);
    double dalpha = MPM_SQRT_TWOTHIRDS * lambdak;
    alpint = s.h0 + dalpha;
    double dfds[6];
#pragma unroll
    for (int c = 0; c < 6; c++) dfds[c] = str[c] / strialmag;
    double dep0 = lambdak * dfds[0], dep1 = lambdak * dfds[1], dep2 = lambdak * dfds[2];
    double dep5 = 2. * lambdak * dfds[5];
    s.alt[0] += dep0; s.alt[1] += dep1; s.alt[2] += dep2; s.alt[5] += dep5;
    double dep3 = 0., dep4 = 0.;
    if (!is2D) {
        dep4 = 2. * lambdak * dfds[4];
        dep3 = 2. * lambdak * dfds[3];
        s.alt[4] += dep4; s.alt[3] += dep3;
    }
    s.sp[0] = str[0] - 2. * m.Gred * dep0;
    s.sp[1] = str[1] - 2. * m.Gred * dep1;
    s.sp[2] = str[2] - 2. * m.Gred * dep2;
    s.sp[5] = str[5] - m.Gred * dep5;
    if (!is2D) {
        s.sp[3] = str[3] - m.Gred * dep3;
        s.sp[4] = str[4] - m.Gred * dep4;
    }
    s.h0 = alpint;
}

// HyperElastic::IncrementDeformation (HyperElastic.cpp:104-141): F <- dF.F, B <- dF.B.dF^T
// with dF = Exponential(1) in 3D and Exponential(2) in 2D; returns det(dF). B is left in s.alt.
template <int DIM>
__device__ __forceinline__ double hyper_increment(const double du[3][3], PState &s) {
    double dF[3][3], pF[3][3], F[3][3], pB[3][3], a[3][3];
    double detDf;
    if (DIM == 3) {
        // Exponential(1) (Matrix3.cpp:358-363)
        dF[0][0] = 1. + du[0][0]; dF[0][1] = du[0][1]; dF[0][2] = du[0][2];
        dF[1][0] = du[1][0]; dF[1][1] = 1. + du[1][1]; dF[1][2] = du[1][2];
        dF[2][0] = du[2][0]; dF[2][1] = du[2][1]; dF[2][2] = 1. + du[2][2];
        get_F3(s.ep, s.wr, pF);
        mat_mul3(dF, pF, F);
        set_F3(F, s.ep, s.wr);
        // GetElasticLeftCauchyMatrix (MatPoint3D.cpp:358-361)
        pB[0][0] = s.alt[0]; pB[0][1] = s.alt[5]; pB[0][2] = s.alt[4];
        pB[1][0] = s.alt[5]; pB[1][1] = s.alt[1]; pB[1][2] = s.alt[3];
        pB[2][0] = s.alt[4]; pB[2][1] = s.alt[3]; pB[2][2] = s.alt[2];
        mat_mul3(dF, pB, a);
        detDf = dF[0][0] * (dF[1][1] * dF[2][2] - dF[2][1] * dF[1][2])
              - dF[1][0] * (dF[0][1] * dF[2][2] - dF[2][1] * dF[0][2])
              + dF[2][0] * (dF[0][1] * dF[1][2] - dF[1][1] * dF[0][2]);
    } else {
        // Exponential(2), 2D branch (Matrix3.cpp:314-347)
        double c0 = du[0][1] * du[1][0] - du[0][0] * du[1][1];
        double c1 = du[0][0] + du[1][1];
        double beta0 = 0., beta1 = 1., alpha0 = 1., alpha1 = 1.;
        double betaz = du[2][2], ezz = 1. + betaz;
        {
            double factor = 1 / (double)2;
            double temp = beta1;
            beta1 = factor * (c1 * temp + beta0);
            beta0 = factor * c0 * temp;
            betaz *= factor * du[2][2];
            alpha0 += beta0;
            alpha1 += beta1;
            ezz += betaz;
        }
        dF[0][0] = alpha0 + alpha1 * du[0][0]; dF[0][1] = alpha1 * du[0][1];
        dF[1][0] = alpha1 * du[1][0]; dF[1][1] = alpha0 + alpha1 * du[1][1]; dF[2][2] = ezz;
        dF[0][2] = dF[1][2] = dF[2][0] = dF[2][1] = 0.;
        pF[0][0] = 1. + s.ep[0]; pF[0][1] = 0.5 * (s.ep[5] - s.wr[2]);
        pF[1][0] = 0.5 * (s.ep[5] + s.wr[2]); pF[1][1] = 1. + s.ep[1]; pF[2][2] = 1. + s.ep[2];
        pF[0][2] = pF[1][2] = pF[2][0] = pF[2][1] = 0.;
        mat_mul2(dF, pF, F);
        s.ep[0] = F[0][0] - 1.; s.ep[1] = F[1][1] - 1.; s.ep[2] = F[2][2] - 1.;
        s.ep[5] = F[1][0] + F[0][1];
        s.wr[2] = F[1][0] - F[0][1];
        pB[0][0] = s.alt[0]; pB[0][1] = s.alt[5]; pB[1][0] = s.alt[5]; pB[1][1] = s.alt[1];
        pB[2][2] = s.alt[2]; pB[0][2] = pB[1][2] = pB[2][0] = pB[2][1] = 0.;
        mat_mul2(dF, pB, a);
        detDf = dF[2][2] * (dF[0][0] * dF[1][1] - dF[1][0] * dF[0][1]);
    }
    double Bxx = a[0][0] * dF[0][0] + a[0][1] * dF[0][1] + a[0][2] * dF[0][2];
    double Bxy = a[0][0] * dF[1][0] + a[0][1] * dF[1][1] + a[0][2] * dF[1][2];
    double Byy = a[1][0] * dF[1][0] + a[1][1] * dF[1][1] + a[1][2] * dF[1][2];
    double Bzz = a[2][0] * dF[2][0] + a[2][1] * dF[2][1] + a[2][2] * dF[2][2];
    s.alt[0] = Bxx; s.alt[5] = Bxy; s.alt[1] = Byy; s.alt[2] = Bzz;
    if (DIM == 3) {
        s.alt[4] = a[0][0] * dF[2][0] + a[0][1] * dF[2][1] + a[0][2] * dF[2][2];
        s.alt[3] = a[1][0] * dF[2][0] + a[1][1] * dF[2][1] + a[1][2] * dF[2][2];
    }
    return detDf;
}

// HyperElastic::GetVolumetricTerms (HyperElastic.cpp:171-207); option numbers from
// HyperElastic.hpp:21
__device__ __forceinline__ double volumetric_term(int ujOption, double J, double Kred) {
    double Kterm;
    if (ujOption == 1) Kterm = Kred * (J - 1.);
    else if (ujOption == 2) { double lj = log(J); Kterm = Kred * lj; Kterm /= J; }
    else Kterm = 0.5 * Kred * (J - 1. / J);
    return Kterm;
}

// Neohookean::MPMConstitutiveLaw (Neohookean.cpp:177-335), no residual strains or artificial
// viscosity; the volumetric term takes the specific Lame constant
template <int DIM>
__device__ __forceinline__ void law_neohookean(const double du[3][3], const MpmMaterial &m, PState &s) {
    double detDf = hyper_increment<DIM>(du, s);
    double Bxx = s.alt[0], Byy = s.alt[1], Bzz = s.alt[2], Bxy = s.alt[5];
    double Bxz = DIM == 3 ? s.alt[4] : 0., Byz = DIM == 3 ? s.alt[3] : 0.;
    double Jres = s.h1;
    double dJres = exp(0.);
    Jres *= dJres;
    s.h1 = Jres;
    double resStretch = pow(Jres, 1. / 3.);
    double Jres23 = resStretch * resStretch;
    double J = detDf * s.h0;
    s.h0 = J;
    double Jeff = J / Jres;
    double Kterm = volumetric_term(m.ujOption, Jeff, m.Lamesp);
    double Pterm = J * Kterm + Jres * m.Gsp * ((Bxx + Byy + Bzz) / (3. * Jres23) - 1.);
    s.P = -Pterm + 0.;
    double GJeff = resStretch * m.Gsp;
    double I1third = (Bxx + Byy + Bzz) / 3.;
    s.sp[0] = GJeff * (Bxx - I1third);
    s.sp[1] = GJeff * (Byy - I1third);
    s.sp[2] = GJeff * (Bzz - I1third);
    s.sp[5] = GJeff * Bxy;
    if (DIM == 3) { s.sp[4] = GJeff * Bxz; s.sp[3] = GJeff * Byz; }
}

// Mooney::MPMConstitutiveLaw (Mooney.cpp:185-367), plane strain or 3D, no residual strains or
// artificial viscosity; the volumetric term takes the specific bulk modulus
template <int DIM>
__device__ __forceinline__ void law_mooney(const double du[3][3], const MpmMaterial &m, PState &s) {
    double detDf = hyper_increment<DIM>(du, s);
    double Bxx = s.alt[0], Byy = s.alt[1], Bzz = s.alt[2], Bxy = s.alt[5];
    double Jres = s.h1;
    double dJres = exp(0.);
    Jres *= dJres;
    s.h1 = Jres;
    double J = detDf * s.h0;
    s.h0 = J;
    double Jeff = J / Jres;
    double Kterm = J * volumetric_term(m.ujOption, Jeff, m.Ksp);
    s.P = -Kterm + 0.;
    double J23 = pow(J, 2. / 3.);
    double J43 = J23 * J23;
    double JforG1 = J23 / Jres;
    double JforG2 = J43 / Jres;
    double sxx = (2 * Bxx - Byy - Bzz) * m.G1sp / (3. * JforG1)
               + (Bxx * (Byy + Bzz) - 2 * Byy * Bzz - Bxy * Bxy) * m.G2sp / (3. * JforG2);
    double syy = (2 * Byy - Bxx - Bzz) * m.G1sp / (3. * JforG1)
               + (Byy * (Bxx + Bzz) - 2 * Bxx * Bzz - Bxy * Bxy) * m.G2sp / (3. * JforG2);
    double szz = (2 * Bzz - Bxx - Byy) * m.G1sp / (3. * JforG1)
               + (Bzz * (Bxx + Byy) - 2 * Bxx * Byy + 2. * Bxy * Bxy) * m.G2sp / (3. * JforG2);
    double sxy = Bxy * m.G1sp / JforG1 + (Bzz * Bxy) * m.G2sp / JforG2;
    if (DIM == 3) {
        double Bxz = s.alt[4], Byz = s.alt[3];
        sxx += (2. * Byz * Byz - Bxz * Bxz) * m.G2sp / (3. * JforG2);
        syy += (2. * Bxz * Bxz - Byz * Byz) * m.G2sp / (3. * JforG2);
        szz -= (Bxz * Bxz + Byz * Byz) * m.G2sp / (3. * JforG2);
        sxy -= Bxz * Byz * m.G2sp / JforG2;
        s.sp[4] = Bxz * m.G1sp / JforG1 + (Byy * Bxz - Bxy * Byz) * m.G2sp / JforG2;
        s.sp[3] = Byz * m.G1sp / JforG1 + (Bxx * Byz - Bxy * Bxz) * m.G2sp / JforG2;
    }
    s.sp[0] = sxx; s.sp[1] = syy; s.sp[2] = szz; s.sp[5] = sxy;
}

// MAT = 0 dispatches on the particle's material type; a nonzero MAT compiles one law only
template <int DIM, int MAT = 0>
__device__ __forceinline__ void constitutive(const double du[3][3], const MpmMaterial &m, PState &s) {
    if (MAT == MPMGPU_MAT_ISOTROPIC) { law_isotropic<DIM>(du, m, s); return; }
    if (MAT == MPMGPU_MAT_ISOPLASTICITY) { law_isoplastic<DIM>(du, m, s); return; }
    if (MAT == MPMGPU_MAT_NEOHOOKEAN) { law_neohookean<DIM>(du, m, s); return; }
    if (MAT == MPMGPU_MAT_MOONEY) { law_mooney<DIM>(du, m, s); return; }
    switch (m.type) {
        case MPMGPU_MAT_ISOTROPIC: law_isotropic<DIM>(du, m, s); break;
        case MPMGPU_MAT_ISOPLASTICITY: law_isoplastic<DIM>(du, m, s); break;
        case MPMGPU_MAT_NEOHOOKEAN: law_neohookean<DIM>(du, m, s); break;
        case MPMGPU_MAT_MOONEY: law_mooney<DIM>(du, m, s); break;
        default: break;
    }
}
