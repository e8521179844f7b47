// Element integrals shared by the host assembly (g++) and the GPU assembly
// (hipcc): one implementation, so both evaluate the same IEEE operations in
// the same order.  The GPU translation unit is compiled without FP
// contraction, which makes its results bit-identical to the host's.
//
// Follows the reference's element library:
//   P1 / P2 shape functions       cpp/libs/Element/src/p1.cpp:35-57, p2.cpp:35-84
//   degree-4 Cowper rule          cpp/libs/Element/src/triangleIntegrator.cpp:27-35
//   local mass / stiffness        triangleIntegrator.cpp:132-190
//   local divergence              triangleIntegrator.cpp:233-293
#pragma once

#if defined(__HIPCC__)
#define NSFEM_HD __host__ __device__
#else
#define NSFEM_HD
#endif

namespace nsfem
{
    namespace elem
    {
        // B = [[m11, -m10], [-m01, m00]] of the affine map x = m xi + b, and
        // det m, so that grad_x = B grad_ref / det.
        struct Geometry
        {
            double b00, b01, b10, b11;
            double det;
        };

        NSFEM_HD inline Geometry geometry(double m00, double m01, double m10, double m11)
        {
            Geometry g;
            g.b00 = m11;
            g.b01 = -m10;
            g.b10 = -m01;
            g.b11 = m00;
            g.det = m00 * m11 - m10 * m01;
            return g;
        }

        // Degree-4, 6-point rule on the reference triangle (0,0),(1,0),(0,1):
        // x, y, weight (weights already include the reference area 1/2).
        NSFEM_HD inline void quadPoint(int q, double & x, double & y, double & w)
        {
            constexpr double a = 9.15762135097710067155318824916321e-02;
            constexpr double b = 8.16847572980457958813360619387822e-01;
            constexpr double c = 4.45948490915965001235576892213430e-01;
            constexpr double d = 1.08103018168069997528846215573139e-01;
            constexpr double wa = 5.49758718276610602870846378209535e-02;
            constexpr double wc = 1.11690794839005624883299105931656e-01;
            switch (q)
            {
            case 0: x = a; y = a; w = wa; break;
            case 1: x = a; y = b; w = wa; break;
            case 2: x = b; y = a; w = wa; break;
            case 3: x = c; y = c; w = wc; break;
            case 4: x = c; y = d; w = wc; break;
            default: x = d; y = c; w = wc; break;
            }
        }

        NSFEM_HD inline void p1Shape(double x, double y, double * phi)
        {
            phi[0] = 1 - x - y;
            phi[1] = x;
            phi[2] = y;
        }

        NSFEM_HD inline void p1Grad(double, double, double * gx, double * gy)
        {
            gx[0] = -1;
            gy[0] = -1;
            gx[1] = 1;
            gy[1] = 0;
            gx[2] = 0;
            gy[2] = 1;
        }

        // Node order: corner (0,0), mid (1/2,0), corner (1,0), mid (1/2,1/2),
        // corner (0,1), mid (0,1/2).
        NSFEM_HD inline void p2Shape(double x, double y, double * phi)
        {
            const double x2 = x * x, y2 = y * y, xy = x * y;
            phi[0] = 1 - 3 * x - 3 * y + 2 * x2 + 2 * y2 + 4 * xy;
            phi[1] = 4 * x - 4 * x2 - 4 * xy;
            phi[2] = -x + 2 * x2;
            phi[3] = 4 * xy;
            phi[4] = -y + 2 * y2;
            phi[5] = 4 * y - 4 * y2 - 4 * xy;
        }

        NSFEM_HD inline void p2Grad(double x, double y, double * gx, double * gy)
        {
            gx[0] = -3 + 4 * x + 4 * y;
            gy[0] = -3 + 4 * y + 4 * x;
            gx[1] = 4 - 8 * x - 4 * y;
            gy[1] = -4 * x;
            gx[2] = -1 + 4 * x;
            gy[2] = 0;
            gx[3] = 4 * y;
            gy[3] = 4 * x;
            gx[4] = 0;
            gy[4] = -1 + 4 * y;
            gx[5] = -4 * y;
            gy[5] = 4 - 8 * y - 4 * x;
        }

        template <int N>
        NSFEM_HD inline void shape(double x, double y, double * phi)
        {
            if constexpr (N == 3)
            {
                p1Shape(x, y, phi);
            }
            else
            {
                p2Shape(x, y, phi);
            }
        }

        template <int N>
        NSFEM_HD inline void grad(double x, double y, double * gx, double * gy)
        {
            if constexpr (N == 3)
            {
                p1Grad(x, y, gx, gy);
            }
            else
            {
                p2Grad(x, y, gx, gy);
            }
        }

        NSFEM_HD inline double absd(double v) { return v < 0 ? -v : v; }

        // Local mass matrix, row-major N x N: int phi_i phi_j.
        template <int N>
        NSFEM_HD inline void mass(const Geometry & g, double * dst)
        {
            const double absDet = absd(g.det);
            double phi[N];
            for (int i = 0; i < N * N; i++)
            {
                dst[i] = 0.0;
            }
            for (int q = 0; q < 6; q++)
            {
                double qx, qy, qw;
                quadPoint(q, qx, qy, qw);
                shape<N>(qx, qy, phi);
                for (int i = 0; i < N; i++)
                {
                    for (int j = 0; j < N; j++)
                    {
                        dst[i * N + j] += phi[i] * phi[j] * qw;
                    }
                }
            }
            for (int i = 0; i < N * N; i++)
            {
                dst[i] *= absDet;
            }
        }

        // Local stiffness matrix, row-major N x N: int grad phi_i . grad phi_j.
        template <int N>
        NSFEM_HD inline void stiffness(const Geometry & g, double * dst)
        {
            const double invAbsDet = 1.0 / absd(g.det);
            const double a00 = g.b00 * g.b00 + g.b10 * g.b10;
            const double a01 = g.b00 * g.b01 + g.b10 * g.b11;
            const double a11 = g.b01 * g.b01 + g.b11 * g.b11;
            double gx[N], gy[N];
            for (int i = 0; i < N * N; i++)
            {
                dst[i] = 0.0;
            }
            for (int q = 0; q < 6; q++)
            {
                double qx, qy, qw;
                quadPoint(q, qx, qy, qw);
                grad<N>(qx, qy, gx, gy);
                const double w = qw * invAbsDet;
                for (int i = 0; i < N; i++)
                {
                    for (int j = 0; j < N; j++)
                    {
                        const double v = gx[i] * (a00 * gx[j] + a01 * gy[j]) + gy[i] * (a01 * gx[j] + a11 * gy[j]);
                        dst[i * N + j] += v * w;
                    }
                }
            }
        }

        // Local divergence blocks, row-major.  Swap = false: rows P1 (3), cols
        // P2 (6), X[r][c] = int d(phi2_c)/dx psi1_r.  Swap = true: rows P2 (6),
        // cols P1 (3), X[r][c] = int d(psi1_c)/dx phi2_r.  Y likewise with d/dy.
        template <bool Swap>
        NSFEM_HD inline void divergence(const Geometry & g, double * dstX, double * dstY)
        {
            constexpr int dofM = Swap ? 3 : 6; // differentiated element
            constexpr int dofS = Swap ? 6 : 3; // test element
            const double sgn = g.det < 0 ? -1.0 : 1.0;
            double phi[dofS], gx[dofM], gy[dofM];
            for (int i = 0; i < dofS * dofM; i++)
            {
                dstX[i] = 0.0;
                dstY[i] = 0.0;
            }
            for (int q = 0; q < 6; q++)
            {
                double qx, qy, qw;
                quadPoint(q, qx, qy, qw);
                shape<dofS>(qx, qy, phi);
                grad<dofM>(qx, qy, gx, gy);
                const double w = qw * sgn;
                for (int r = 0; r < dofS; r++)
                {
                    for (int c = 0; c < dofM; c++)
                    {
                        const double ggx = g.b00 * gx[c] + g.b01 * gy[c];
                        const double ggy = g.b10 * gx[c] + g.b11 * gy[c];
                        dstX[r * dofM + c] += ggx * phi[r] * w;
                        dstY[r * dofM + c] += ggy * phi[r] * w;
                    }
                }
            }
        }
    } // namespace elem
} // namespace nsfem
