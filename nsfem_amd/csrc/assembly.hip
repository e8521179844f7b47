#include "assembly.hpp"

#include <stdexcept>

#include "common.hpp"
#include "element_math.hpp"
#include "ops.hpp"

namespace nsfem
{
    namespace
    {
        // Local matrices of element e, element-major: mass and stiffness (P2,
        // 36 each), pressure stiffness (P1, 9), VPD (P1 x P2, x then y, 36),
        // PVD (P2 x P1, x then y, 36).
        __global__ void elementMatricesKernel(int E, const double * __restrict__ affineV, const double * __restrict__ affineP,
                                              double * __restrict__ massLoc, double * __restrict__ stiffLoc,
                                              double * __restrict__ pStiffLoc, double * __restrict__ vpdLoc,
                                              double * __restrict__ pvdLoc)
        {
            const int e = blockIdx.x * blockDim.x + threadIdx.x;
            if (e >= E)
            {
                return;
            }
            const size_t e4 = static_cast<size_t>(e) * 4;
            const elem::Geometry gv = elem::geometry(affineV[e4], affineV[e4 + 1], affineV[e4 + 2], affineV[e4 + 3]);
            const elem::Geometry gp = elem::geometry(affineP[e4], affineP[e4 + 1], affineP[e4 + 2], affineP[e4 + 3]);
            const size_t e36 = static_cast<size_t>(e) * 36;
            double loc[36];
            elem::mass<6>(gv, loc);
            for (int k = 0; k < 36; k++)
            {
                massLoc[e36 + k] = loc[k];
            }
            elem::stiffness<6>(gv, loc);
            for (int k = 0; k < 36; k++)
            {
                stiffLoc[e36 + k] = loc[k];
            }
            elem::stiffness<3>(gp, loc);
            for (int k = 0; k < 9; k++)
            {
                pStiffLoc[static_cast<size_t>(e) * 9 + k] = loc[k];
            }
            double lx[18], ly[18];
            elem::divergence<false>(gv, lx, ly);
            for (int k = 0; k < 18; k++)
            {
                vpdLoc[e36 + k] = lx[k];
                vpdLoc[e36 + 18 + k] = ly[k];
            }
            elem::divergence<true>(gv, lx, ly);
            for (int k = 0; k < 18; k++)
            {
                pvdLoc[e36 + k] = lx[k];
                pvdLoc[e36 + 18 + k] = ly[k];
            }
        }

        // values[i] = sum of loc over entry i's slots, in list order (the
        // first term starts the sum, as in the host assembly).
        __global__ void gatherSlotsKernel(int nnz, const int * __restrict__ start, const int * __restrict__ slot,
                                          const double * __restrict__ loc, double * __restrict__ values)
        {
            const int i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= nnz)
            {
                return;
            }
            const int k0 = start[i], k1 = start[i + 1];
            double acc = loc[slot[k0]];
            for (int k = k0 + 1; k < k1; k++)
            {
                acc += loc[slot[k]];
            }
            values[i] = acc;
        }

        std::vector<double> affineOf(const ConcreteMesh & m)
        {
            std::vector<double> a(static_cast<size_t>(m.numElements) * 4);
            for (int e = 0; e < m.numElements; e++)
            {
                const auto & t = m.elementTransforms[e];
                a[static_cast<size_t>(e) * 4] = t.m[0][0];
                a[static_cast<size_t>(e) * 4 + 1] = t.m[0][1];
                a[static_cast<size_t>(e) * 4 + 2] = t.m[1][0];
                a[static_cast<size_t>(e) * 4 + 3] = t.m[1][1];
            }
            return a;
        }

        void gather(const CsrD & pattern, const SlotList & slots, const DeviceArray<double> & loc, CsrD & out, hipStream_t s)
        {
            out.rows = pattern.rows;
            out.cols = pattern.cols;
            out.rowStart = pattern.rowStart;
            out.column = pattern.column;
            const int nnz = pattern.nnz();
            out.values.assign(nnz, 0.0);
            if (nnz == 0)
            {
                return;
            }
            for (int i = 0; i < nnz; i++)
            {
                if (slots.start[i + 1] <= slots.start[i])
                {
                    throw std::runtime_error("assembly: a pattern entry without contributions");
                }
            }
            DeviceArray<int> start, slot;
            DeviceArray<double> values;
            start.upload(slots.start);
            slot.upload(slots.slot);
            values.resize(nnz);
            gatherSlotsKernel<<<(nnz + 255) / 256, 256, 0, s>>>(nnz, start.get(), slot.get(), loc.get(), values.get());
            NSFEM_HIP(hipGetLastError());
            NSFEM_HIP(hipStreamSynchronize(s));
            out.values = values.download();
        }
    } // namespace

    ChorinMatrices assembleChorinMatricesGpu(const ConcreteMesh & vm, const ConcreteMesh & pm, const AssemblyTopology & topo)
    {
        const int E = vm.numElements;
        if (pm.numElements != E)
        {
            throw std::invalid_argument("Velocity and pressure meshes have different number of elements");
        }
        ChorinMatrices out;
        if (E == 0)
        {
            throw std::invalid_argument("empty mesh");
        }
        const hipStream_t s = libraryStream();
        DeviceArray<double> affV, affP, massLoc, stiffLoc, pStiffLoc, vpdLoc, pvdLoc;
        affV.upload(affineOf(vm));
        affP.upload(affineOf(pm));
        const size_t n36 = static_cast<size_t>(E) * 36;
        massLoc.resize(n36);
        stiffLoc.resize(n36);
        pStiffLoc.resize(static_cast<size_t>(E) * 9);
        vpdLoc.resize(n36);
        pvdLoc.resize(n36);
        elementMatricesKernel<<<(E + 127) / 128, 128, 0, s>>>(E, affV.get(), affP.get(), massLoc.get(), stiffLoc.get(), pStiffLoc.get(),
                                                            vpdLoc.get(), pvdLoc.get());
        NSFEM_HIP(hipGetLastError());
        gather(topo.velocity, topo.velocitySlots, massLoc, out.velocityMass, s);
        gather(topo.velocity, topo.velocitySlots, stiffLoc, out.velocityStiffness, s);
        gather(topo.pressure, topo.pressureSlots, pStiffLoc, out.pressureStiffness, s);
        gather(topo.velocityPressureDiv, topo.vpdSlots, vpdLoc, out.velocityPressureDiv, s);
        gather(topo.pressureVelocityDiv, topo.pvdSlots, pvdLoc, out.pressureVelocityDiv, s);
        return out;
    }
} // namespace nsfem
