#include "fem.hpp"

#include "element_math.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <stdexcept>
#include <thread>

namespace nsfem
{
    int defaultThreads()
    {
        unsigned n = std::thread::hardware_concurrency();
        if (n == 0)
        {
            n = 4;
        }
        return static_cast<int>(std::min(n, 16u));
    }

    const std::vector<QuadPoint> & quadratureDeg4()
    {
        // The published 6-point degree-4 rule (Cowper 1973 / Strang-Fix), with
        // weights already multiplied by the reference-triangle area 1/2.
        static const std::vector<QuadPoint> pts = {
            {9.15762135097710067155318824916321e-02, 9.15762135097710067155318824916321e-02, 5.49758718276610602870846378209535e-02},
            {9.15762135097710067155318824916321e-02, 8.16847572980457958813360619387822e-01, 5.49758718276610602870846378209535e-02},
            {8.16847572980457958813360619387822e-01, 9.15762135097710067155318824916321e-02, 5.49758718276610602870846378209535e-02},
            {4.45948490915965001235576892213430e-01, 4.45948490915965001235576892213430e-01, 1.11690794839005624883299105931656e-01},
            {4.45948490915965001235576892213430e-01, 1.08103018168069997528846215573139e-01, 1.11690794839005624883299105931656e-01},
            {1.08103018168069997528846215573139e-01, 4.45948490915965001235576892213430e-01, 1.11690794839005624883299105931656e-01},
        };
        return pts;
    }

    void p1Shape(double x, double y, double * phi)
    {
        phi[0] = 1 - x - y;
        phi[1] = x;
        phi[2] = y;
    }

    void p1Grad(double, double, double * gx, double * gy)
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
    void p2Shape(double x, double y, double * phi)
    {
        const double x2 = x * x, y2 = y * y, xy = x * y;
        phi[0] = 1 - 3 * x - 3 * y + 2 * x2 + 2 * y2 + 4 * xy;
        phi[1] = 4 * x - 4 * x2 - 4 * xy;
        phi[2] = -x + 2 * x2;
        phi[3] = 4 * xy;
        phi[4] = -y + 2 * y2;
        phi[5] = 4 * y - 4 * y2 - 4 * xy;
    }

    void p2Grad(double x, double y, double * gx, double * gy)
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

    namespace
    {
        void shapeOf(int nDof, double x, double y, double * phi)
        {
            if (nDof == 3)
            {
                p1Shape(x, y, phi);
            }
            else
            {
                p2Shape(x, y, phi);
            }
        }

        void gradOf(int nDof, double x, double y, double * gx, double * gy)
        {
            if (nDof == 3)
            {
                p1Grad(x, y, gx, gy);
            }
            else
            {
                p2Grad(x, y, gx, gy);
            }
        }
    } // namespace

    ElementGeometry geometryOf(const AffineMap & t)
    {
        ElementGeometry g;
        g.b00 = t.m[1][1];
        g.b01 = -t.m[1][0];
        g.b10 = -t.m[0][1];
        g.b11 = t.m[0][0];
        g.det = t.m[0][0] * t.m[1][1] - t.m[1][0] * t.m[0][1];
        return g;
    }

    namespace
    {
        elem::Geometry elementGeometry(const AffineMap & t)
        {
            return elem::geometry(t.m[0][0], t.m[0][1], t.m[1][0], t.m[1][1]);
        }
    } // namespace

    void localMass(const AffineMap & t, int nDof, double * dst)
    {
        if (nDof == 3)
        {
            elem::mass<3>(elementGeometry(t), dst);
        }
        else
        {
            elem::mass<6>(elementGeometry(t), dst);
        }
    }

    void localStiffness(const AffineMap & t, int nDof, double * dst)
    {
        if (nDof == 3)
        {
            elem::stiffness<3>(elementGeometry(t), dst);
        }
        else
        {
            elem::stiffness<6>(elementGeometry(t), dst);
        }
    }

    void localSelfConvectionP2(const AffineMap & t, const double * ux, const double * uy, double * dst)
    {
        const ElementGeometry g = geometryOf(t);
        const double sgn = g.det < 0 ? -1.0 : 1.0;
        double phi[6], gx[6], gy[6];
        std::fill(dst, dst + 36, 0.0);
        for (const auto & q : quadratureDeg4())
        {
            p2Shape(q.x, q.y, phi);
            p2Grad(q.x, q.y, gx, gy);
            double fx = 0, fy = 0;
            for (int i = 0; i < 6; i++)
            {
                fx += phi[i] * ux[i];
                fy += phi[i] * uy[i];
            }
            // Flow in reference coordinates: B^T u
            const double lx = g.b00 * fx + g.b10 * fy;
            const double ly = g.b01 * fx + g.b11 * fy;
            const double w = q.w * sgn;
            for (int r = 0; r < 6; r++)
            {
                for (int c = 0; c < 6; c++)
                {
                    dst[r * 6 + c] += (lx * gx[c] + ly * gy[c]) * phi[r] * w;
                }
            }
        }
    }

    void localDivergence(const AffineMap & t, bool swap, double * dstX, double * dstY)
    {
        if (swap)
        {
            elem::divergence<true>(elementGeometry(t), dstX, dstY);
        }
        else
        {
            elem::divergence<false>(elementGeometry(t), dstX, dstY);
        }
    }

    const ConvectionTensors & convectionTensors()
    {
        static const ConvectionTensors T = []
        {
            ConvectionTensors t{};
            double phi[6], gx[6], gy[6];
            for (const auto & q : quadratureDeg4())
            {
                p2Shape(q.x, q.y, phi);
                p2Grad(q.x, q.y, gx, gy);
                for (int r = 0; r < 6; r++)
                {
                    for (int c = 0; c < 6; c++)
                    {
                        for (int k = 0; k < 6; k++)
                        {
                            t.tx[r][c][k] += q.w * phi[r] * phi[k] * gx[c];
                            t.ty[r][c][k] += q.w * phi[r] * phi[k] * gy[c];
                        }
                    }
                }
            }
            return t;
        }();
        return T;
    }

    namespace
    {
        // node -> list of (element, local index), in ascending element order.
        struct Incidence
        {
            std::vector<int> start;
            std::vector<int> element;
            std::vector<int> local;
        };

        Incidence buildIncidence(const ConcreteMesh & mesh)
        {
            const int n = static_cast<int>(mesh.nodes.size());
            const int npe = mesh.nodesPerElement;
            Incidence inc;
            inc.start.assign(n + 1, 0);
            for (int id : mesh.elements)
            {
                inc.start[id + 1]++;
            }
            for (int i = 0; i < n; i++)
            {
                inc.start[i + 1] += inc.start[i];
            }
            inc.element.resize(mesh.elements.size());
            inc.local.resize(mesh.elements.size());
            std::vector<int> fill(inc.start.begin(), inc.start.end() - 1);
            for (int e = 0; e < mesh.numElements; e++)
            {
                for (int a = 0; a < npe; a++)
                {
                    const int id = mesh.elements[static_cast<size_t>(e) * npe + a];
                    const int pos = fill[id]++;
                    inc.element[pos] = e;
                    inc.local[pos] = a;
                }
            }
            return inc;
        }

        // Row-parallel CSR builder.  rowFn(r, contributions) appends (col, val)
        // pairs in a fixed order; duplicates are summed in that order, so the
        // result is independent of the thread count.
        using RowFn = std::function<void(int, std::vector<std::pair<int, double>> &)>;

        CsrD assembleRows(int rows, int cols, int nThreads, const RowFn & rowFn)
        {
            nThreads = std::max(1, std::min(nThreads, rows));
            struct Part
            {
                std::vector<int> counts;
                std::vector<int> column;
                std::vector<double> values;
            };
            std::vector<Part> parts(nThreads);
            auto work = [&](int t)
            {
                const int r0 = static_cast<int>(static_cast<long long>(rows) * t / nThreads);
                const int r1 = static_cast<int>(static_cast<long long>(rows) * (t + 1) / nThreads);
                Part & p = parts[t];
                p.counts.reserve(r1 - r0);
                std::vector<std::pair<int, double>> contrib;
                for (int r = r0; r < r1; r++)
                {
                    contrib.clear();
                    rowFn(r, contrib);
                    std::stable_sort(contrib.begin(), contrib.end(),
                                     [](const auto & a, const auto & b) { return a.first < b.first; });
                    int count = 0;
                    for (size_t i = 0; i < contrib.size(); i++)
                    {
                        if (i > 0 && contrib[i].first == contrib[i - 1].first)
                        {
                            p.values.back() += contrib[i].second;
                        }
                        else
                        {
                            if (contrib[i].first < 0 || contrib[i].first >= cols)
                            {
                                throw std::runtime_error("assembleRows: column out of range");
                            }
                            p.column.push_back(contrib[i].first);
                            p.values.push_back(contrib[i].second);
                            count++;
                        }
                    }
                    p.counts.push_back(count);
                }
            };
            std::vector<std::thread> threads;
            std::vector<std::exception_ptr> errors(nThreads);
            for (int t = 0; t < nThreads; t++)
            {
                threads.emplace_back([&, t]
                                     {
                    try { work(t); }
                    catch (...) { errors[t] = std::current_exception(); } });
            }
            for (auto & th : threads)
            {
                th.join();
            }
            for (auto & e : errors)
            {
                if (e)
                {
                    std::rethrow_exception(e);
                }
            }
            CsrD m;
            m.rows = rows;
            m.cols = cols;
            m.rowStart.resize(rows + 1);
            size_t nnz = 0;
            for (auto & p : parts)
            {
                nnz += p.column.size();
            }
            if (nnz > 0x7fffffffULL)
            {
                throw std::runtime_error("assembleRows: more than 2^31 non-zeros");
            }
            m.column.reserve(nnz);
            m.values.reserve(nnz);
            int r = 0;
            for (auto & p : parts)
            {
                for (int c : p.counts)
                {
                    m.rowStart[r + 1] = m.rowStart[r] + c;
                    r++;
                }
                m.column.insert(m.column.end(), p.column.begin(), p.column.end());
                m.values.insert(m.values.end(), p.values.begin(), p.values.end());
                p = Part{};
            }
            return m;
        }
    } // namespace

    ChorinMatrices buildChorinMatrices(const ConcreteMesh & vm, const ConcreteMesh & pm, int nThreads)
    {
        if (vm.numElements != pm.numElements)
        {
            throw std::invalid_argument("Velocity and pressure meshes have different number of elements");
        }
        if (vm.type != ElementType::P2 || pm.type != ElementType::P1)
        {
            throw std::invalid_argument("Chorin matrices need a P2 velocity mesh and a P1 pressure mesh");
        }
        const int N = static_cast<int>(vm.nodes.size());
        const int V = static_cast<int>(pm.nodes.size());
        const Incidence incV = buildIncidence(vm);
        const Incidence incP = buildIncidence(pm);

        ChorinMatrices out;
        // Velocity mass and stiffness share the P2 pattern.
        {
            CsrD mass = assembleRows(N, N, nThreads, [&](int r, auto & c)
                                     {
                double loc[36];
                for (int k = incV.start[r]; k < incV.start[r + 1]; k++)
                {
                    const int e = incV.element[k], a = incV.local[k];
                    localMass(vm.elementTransforms[e], 6, loc);
                    const int * ids = vm.elements.data() + static_cast<size_t>(e) * 6;
                    for (int b = 0; b < 6; b++) c.emplace_back(ids[b], loc[a * 6 + b]);
                } });
            CsrD stiff = assembleRows(N, N, nThreads, [&](int r, auto & c)
                                      {
                double loc[36];
                for (int k = incV.start[r]; k < incV.start[r + 1]; k++)
                {
                    const int e = incV.element[k], a = incV.local[k];
                    localStiffness(vm.elementTransforms[e], 6, loc);
                    const int * ids = vm.elements.data() + static_cast<size_t>(e) * 6;
                    for (int b = 0; b < 6; b++) c.emplace_back(ids[b], loc[a * 6 + b]);
                } });
            out.velocityMass = std::move(mass);
            out.velocityStiffness = std::move(stiff);
        }
        out.pressureStiffness = assembleRows(V, V, nThreads, [&](int r, auto & c)
                                             {
            double loc[9];
            for (int k = incP.start[r]; k < incP.start[r + 1]; k++)
            {
                const int e = incP.element[k], a = incP.local[k];
                localStiffness(pm.elementTransforms[e], 3, loc);
                const int * ids = pm.elements.data() + static_cast<size_t>(e) * 3;
                for (int b = 0; b < 3; b++) c.emplace_back(ids[b], loc[a * 3 + b]);
            } });
        // VPD: row = pressure node p, cols gV (x) and gV + N (y)
        out.velocityPressureDiv = assembleRows(V, 2 * N, nThreads, [&](int r, auto & c)
                                               {
            double lx[18], ly[18];
            for (int k = incP.start[r]; k < incP.start[r + 1]; k++)
            {
                const int e = incP.element[k], a = incP.local[k];
                localDivergence(vm.elementTransforms[e], false, lx, ly);
                const int * ids = vm.elements.data() + static_cast<size_t>(e) * 6;
                for (int b = 0; b < 6; b++)
                {
                    c.emplace_back(ids[b], lx[a * 6 + b]);
                    c.emplace_back(ids[b] + N, ly[a * 6 + b]);
                }
            } });
        // PVD: rows v (x) and v + N (y), cols = pressure nodes
        out.pressureVelocityDiv = assembleRows(2 * N, V, nThreads, [&](int row, auto & c)
                                               {
            const bool isY = row >= N;
            const int r = isY ? row - N : row;
            double lx[18], ly[18];
            for (int k = incV.start[r]; k < incV.start[r + 1]; k++)
            {
                const int e = incV.element[k], a = incV.local[k];
                localDivergence(vm.elementTransforms[e], true, lx, ly);
                const int * ids = pm.elements.data() + static_cast<size_t>(e) * 3;
                const double * l = isY ? ly : lx;
                for (int b = 0; b < 3; b++) c.emplace_back(ids[b], l[a * 3 + b]);
            } });
        return out;
    }

    CsrD buildFastConvectionMap(const ConcreteMesh & vm, const CsrD & pattern, int nThreads)
    {
        const int N = static_cast<int>(vm.nodes.size());
        if (pattern.rows != N || pattern.cols != N)
        {
            throw std::invalid_argument("buildFastConvectionMap: pattern must be N x N");
        }
        const Incidence inc = buildIncidence(vm);
        // Local convection is linear in the 12 element velocities; probe each
        // one with a unit value (fastConvection.cpp:69-91).
        const int E = vm.numElements;
        // rowOf[k] = CSR entry index of the k-th (r, c) pair, for a given element.
        const long long nnzC = pattern.nnz();
        if (nnzC > 0x7fffffffLL)
        {
            throw std::runtime_error("FastConvection: too many entries");
        }
        // Build row by row over entries of C: entry (r, c) collects its
        // contributions from every element that contains both r and c.
        return assembleRows(static_cast<int>(nnzC), 2 * N, nThreads, [&](int entry, auto & contrib)
                            {
            // Find the row r of this entry.
            const auto it = std::upper_bound(pattern.rowStart.begin(), pattern.rowStart.end(), entry);
            const int r = static_cast<int>(it - pattern.rowStart.begin()) - 1;
            const int col = pattern.column[entry];
            double ux[6] = {0}, uy[6] = {0}, loc[36];
            for (int k = inc.start[r]; k < inc.start[r + 1]; k++)
            {
                const int e = inc.element[k], a = inc.local[k];
                const int * ids = vm.elements.data() + static_cast<size_t>(e) * 6;
                int b = -1;
                for (int j = 0; j < 6; j++) if (ids[j] == col) { b = j; break; }
                if (b < 0) continue;
                for (int p = 0; p < 12; p++)
                {
                    const bool isY = p >= 6;
                    const int l = isY ? p - 6 : p;
                    (isY ? uy : ux)[l] = 1;
                    localSelfConvectionP2(vm.elementTransforms[e], ux, uy, loc);
                    (isY ? uy : ux)[l] = 0;
                    contrib.emplace_back(isY ? ids[l] + N : ids[l], loc[a * 6 + b]);
                }
            }
            (void)E; });
    }

    namespace
    {
        // Row-parallel pattern + slot-list builder: rowFn(r, pairs) appends
        // (column, slot) pairs; a row's entries are its distinct columns and
        // each entry's slots come out in ascending slot order (elements are
        // visited in ascending order, one slot per element and entry).
        using SlotRowFn = std::function<void(int, std::vector<std::pair<int, int>> &)>;

        void buildRowsWithSlots(int rows, int cols, int nThreads, const SlotRowFn & rowFn, CsrD & pattern, SlotList & slots)
        {
            nThreads = std::max(1, std::min(nThreads, rows));
            struct Part
            {
                std::vector<int> rowCounts, column, entryCounts, slot;
            };
            std::vector<Part> parts(nThreads);
            std::vector<std::exception_ptr> errors(nThreads);
            std::vector<std::thread> threads;
            for (int t = 0; t < nThreads; t++)
            {
                threads.emplace_back([&, t]
                                     {
                    try
                    {
                        const int r0 = static_cast<int>(static_cast<long long>(rows) * t / nThreads);
                        const int r1 = static_cast<int>(static_cast<long long>(rows) * (t + 1) / nThreads);
                        Part & p = parts[t];
                        std::vector<std::pair<int, int>> pairs;
                        for (int r = r0; r < r1; r++)
                        {
                            pairs.clear();
                            rowFn(r, pairs);
                            std::sort(pairs.begin(), pairs.end());
                            int count = 0;
                            for (size_t i = 0; i < pairs.size(); i++)
                            {
                                if (pairs[i].first < 0 || pairs[i].first >= cols)
                                {
                                    throw std::runtime_error("assembly: column out of range");
                                }
                                if (i == 0 || pairs[i].first != pairs[i - 1].first)
                                {
                                    p.column.push_back(pairs[i].first);
                                    p.entryCounts.push_back(0);
                                    count++;
                                }
                                p.entryCounts.back()++;
                                p.slot.push_back(pairs[i].second);
                            }
                            p.rowCounts.push_back(count);
                        }
                    }
                    catch (...)
                    {
                        errors[t] = std::current_exception();
                    } });
            }
            for (auto & th : threads)
            {
                th.join();
            }
            for (auto & e : errors)
            {
                if (e)
                {
                    std::rethrow_exception(e);
                }
            }
            pattern = CsrD{};
            pattern.rows = rows;
            pattern.cols = cols;
            pattern.rowStart.assign(rows + 1, 0);
            slots.start.assign(1, 0);
            int r = 0;
            for (auto & p : parts)
            {
                for (int c : p.rowCounts)
                {
                    pattern.rowStart[r + 1] = pattern.rowStart[r] + c;
                    r++;
                }
                pattern.column.insert(pattern.column.end(), p.column.begin(), p.column.end());
                for (int c : p.entryCounts)
                {
                    slots.start.push_back(slots.start.back() + c);
                }
                slots.slot.insert(slots.slot.end(), p.slot.begin(), p.slot.end());
                p = Part{};
            }
        }
    } // namespace

    AssemblyTopology buildAssemblyTopology(const ConcreteMesh & vm, const ConcreteMesh & pm, int nThreads)
    {
        if (vm.numElements != pm.numElements || vm.type != ElementType::P2 || pm.type != ElementType::P1)
        {
            throw std::invalid_argument("Assembly needs a P2 velocity mesh and a P1 pressure mesh over the same elements");
        }
        const int N = static_cast<int>(vm.nodes.size());
        const int V = static_cast<int>(pm.nodes.size());
        const int E = vm.numElements;
        if (static_cast<long long>(E) * 36 > 0x7fffffffLL || 2LL * N > 0x7fffffffLL)
        {
            throw std::invalid_argument("mesh too large for 32-bit slot indices");
        }
        const Incidence incV = buildIncidence(vm);
        const Incidence incP = buildIncidence(pm);
        const int * ev = vm.elements.data();
        const int * ep = pm.elements.data();
        AssemblyTopology t;
        // velocity (M, K): row r, element e with r at local a: (ids6[b], e*36 + a*6 + b)
        buildRowsWithSlots(N, N, nThreads, [&](int r, auto & pr)
                           {
            for (int k = incV.start[r]; k < incV.start[r + 1]; k++)
            {
                const int e = incV.element[k], a = incV.local[k];
                for (int b = 0; b < 6; b++) pr.emplace_back(ev[static_cast<size_t>(e) * 6 + b], e * 36 + a * 6 + b);
            } }, t.velocity, t.velocitySlots);
        // pressure stiffness: (ids3[b], e*9 + a*3 + b)
        buildRowsWithSlots(V, V, nThreads, [&](int r, auto & pr)
                           {
            for (int k = incP.start[r]; k < incP.start[r + 1]; k++)
            {
                const int e = incP.element[k], a = incP.local[k];
                for (int b = 0; b < 3; b++) pr.emplace_back(ep[static_cast<size_t>(e) * 3 + b], e * 9 + a * 3 + b);
            } }, t.pressure, t.pressureSlots);
        // VPD: pressure row, (ids6[b] + comp N, e*36 + comp*18 + a*6 + b)
        buildRowsWithSlots(V, 2 * N, nThreads, [&](int r, auto & pr)
                           {
            for (int k = incP.start[r]; k < incP.start[r + 1]; k++)
            {
                const int e = incP.element[k], a = incP.local[k];
                for (int b = 0; b < 6; b++)
                {
                    const int id = ev[static_cast<size_t>(e) * 6 + b];
                    pr.emplace_back(id, e * 36 + a * 6 + b);
                    pr.emplace_back(id + N, e * 36 + 18 + a * 6 + b);
                }
            } }, t.velocityPressureDiv, t.vpdSlots);
        // PVD: velocity row (x, then y + N), (ids3[b], e*36 + comp*18 + a*3 + b)
        buildRowsWithSlots(2 * N, V, nThreads, [&](int row, auto & pr)
                           {
            const int comp = row >= N ? 1 : 0;
            const int r = row - comp * N;
            for (int k = incV.start[r]; k < incV.start[r + 1]; k++)
            {
                const int e = incV.element[k], a = incV.local[k];
                for (int b = 0; b < 3; b++) pr.emplace_back(ep[static_cast<size_t>(e) * 3 + b], e * 36 + comp * 18 + a * 3 + b);
            } }, t.pressureVelocityDiv, t.pvdSlots);
        return t;
    }

    ConvectionElementData buildConvectionElementData(const ConcreteMesh & vm, const CsrD & pattern, const SlotList * slots)
    {
        ConvectionElementData d;
        const int E = vm.numElements;
        d.numElements = E;
        d.elementNodes = vm.elements;
        d.geometry.resize(static_cast<size_t>(E) * 5);
        for (int e = 0; e < E; e++)
        {
            const ElementGeometry g = geometryOf(vm.elementTransforms[e]);
            double * dst = d.geometry.data() + static_cast<size_t>(e) * 5;
            dst[0] = g.b00;
            dst[1] = g.b01;
            dst[2] = g.b10;
            dst[3] = g.b11;
            dst[4] = g.det < 0 ? -1.0 : 1.0;
        }
        const int nnz = pattern.nnz();
        if (slots)
        {
            if (static_cast<int>(slots->start.size()) != nnz + 1)
            {
                throw std::invalid_argument("slot lists do not match the velocity pattern");
            }
            d.contribStart = slots->start;
            d.contribIndex = slots->slot;
            return d;
        }
        d.contribStart.assign(nnz + 1, 0);
        std::vector<int> slot(static_cast<size_t>(E) * 36);
        for (int e = 0; e < E; e++)
        {
            const int * ids = vm.elements.data() + static_cast<size_t>(e) * 6;
            for (int r = 0; r < 6; r++)
            {
                for (int c = 0; c < 6; c++)
                {
                    const int pos = pattern.find(ids[r], ids[c]);
                    if (pos < 0)
                    {
                        throw std::runtime_error("Convection slot missing from the velocity pattern");
                    }
                    slot[static_cast<size_t>(e) * 36 + r * 6 + c] = pos;
                    d.contribStart[pos + 1]++;
                }
            }
        }
        for (int i = 0; i < nnz; i++)
        {
            d.contribStart[i + 1] += d.contribStart[i];
        }
        d.contribIndex.resize(static_cast<size_t>(E) * 36);
        std::vector<int> fill(d.contribStart.begin(), d.contribStart.end() - 1);
        for (size_t k = 0; k < slot.size(); k++)
        {
            d.contribIndex[fill[slot[k]]++] = static_cast<int>(k);
        }
        return d;
    }
} // namespace nsfem
