// Exercises the C++ host layer (include/msolve_amd.hpp) end to end on the
// GPU: a plain SolverPCG solve (Jacobi and multigrid), a two-class Monte Carlo batch, a Newmark
// run and SolverSkyline's zero-energy modes on a floating box, printing
// the results as JSON for tests/test_gpu_host_cpp.py to compare with the
// oracle.  Usage: host_api_check <kl_table.txt> (M, then rows of w a per
// axis, evpPos, eigs) ; xi and moduli are derived deterministically.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "../../include/msolve_amd.hpp"

static void dump(const char *key, const std::vector<double> &v, bool last = false)
{
    std::printf("\"%s\": [", key);
    for (size_t i = 0; i < v.size(); i++) std::printf(i ? ",%.17g" : "%.17g", v[i]);
    std::printf("]%s\n", last ? "" : ",");
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: host_api_check kl_table.txt\n");
        return 2;
    }
    try {
        std::ifstream in(argv[1]);
        msolve::KLTable t;
        in >> t.M >> t.kmax;
        t.w1d.resize(3 * t.kmax);
        t.a1d.resize(3 * t.kmax);
        t.evpPos.resize(3 * t.M);
        t.eigs.resize(t.M);
        for (auto &v : t.w1d) in >> v;
        for (auto &v : t.a1d) in >> v;
        for (auto &v : t.evpPos) in >> v;
        for (auto &v : t.eigs) in >> v;
        std::vector<double> xi;
        double v;
        while (in >> v) xi.push_back(v);

        msolve::Context ctx(0);
        msolve::Model m6 = msolve::Hexa8MeshGenerator::Cantilever(6);
        std::vector<double> E(m6.ElementCount(), 2.1e5);
        msolve::SolverPCG solver(ctx, m6, E, 0.3);
        solver.Initialize();
        solver.Solve(20000, 1e-12);

        msolve::Model m4 = msolve::Hexa8MeshGenerator::Cantilever(4);
        msolve::MonteCarloAnalyzerWithStochasticMaterial mc(ctx, {&m4, &m6}, {3, 2}, t, xi, 2.1e5,
                                                            0.3, 0.3, MSA_FP64);
        mc.Initialize();
        mc.Solve(1e-12, 20000, 1e-4, 30, 16);

        msolve::ElementStructuralStiffnessProvider prov(ctx);
        std::vector<double> E4(m4.ElementCount(), 1.0);
        std::vector<double> ke = prov.Matrix(m4, E4, 0.3);

        std::printf("{\n\"status\": %d, \"current_iteration\": %d, \"products\": %d,\n",
                    solver.Status, solver.CurrentIteration, solver.Products);
        dump("u_plain", solver.VectorX);
        std::vector<int> st4 = mc.Status(0), st6 = mc.Status(1);
        int bad = 0;
        for (int s : st4) bad += s != MSA_OK;
        for (int s : st6) bad += s != MSA_OK;
        std::printf("\"mc_bad\": %d,\n", bad);
        for (int s = 0; s < 3; s++) {
            std::string k = "u4_" + std::to_string(s);
            dump(k.c_str(), mc.Solution(0, s));
        }
        for (int s = 0; s < 2; s++) {
            std::string k = "u6_" + std::to_string(s);
            dump(k.c_str(), mc.Solution(1, s));
        }
        // Newmark of the 4^3 block: density 50, dt 2e-3, 30 steps of a 20 Hz
        // base acceleration in Z plus the static tip load, FP64 PCG to 1e-12
        std::vector<double> E4n(m4.ElementCount(), 2.1e5);
        msolve::MassAccelerationSinusoidalLoad acc(0.05, 2e-3, 30 * 2e-3, 1.0);
        msolve::NewmarkDynamicAnalyzer nm(ctx, m4, E4n, 1, 0.25, 0.5, 2e-3, 30 * 2e-3, 50.0, acc,
                                          MSA_FP64);
        nm.BuildMatrices();
        nm.Solve(1e-12, 20000, 8);
        dump("newmark_hist", nm.History(0));
        // the same SolverPCG surface with the multigrid preconditioner (8^3)
        msolve::Model m8 = msolve::Hexa8MeshGenerator::Cantilever(8);
        std::vector<double> E8(m8.ElementCount(), 2.1e5);
        msolve::SolverPCG mg(ctx, m8, E8, 0.3);
        mg.Preconditioner = MSA_PRECOND_MULTIGRID;
        mg.Initialize();
        mg.Solve(20000, 1e-12);
        std::printf("\"mg_products\": %d, \"mg_status\": %d,\n", mg.Products, mg.Status);
        dump("u_mg8", mg.VectorX);
        // SolverSkyline with zero-energy modes on a floating 3^3 box (no
        // supports: six rigid-body modes), unit load on the last DOF
        msolve::Model fl = msolve::Hexa8MeshGenerator::Box(3, 3, 3);
        fl.Loads.push_back({fl.NodeCount() - 1, msolve::DOFType::X, 1.0});
        fl.ConnectDataStructures();
        std::vector<double> Efl(fl.ElementCount(), 2.1e5);
        msolve::SolverSkyline sky(ctx, fl, Efl, 0.3);
        std::vector<std::vector<double>> zems;
        std::vector<int> zemCols;
        sky.Factorize(1e-3, zems, zemCols);
        sky.Solve();
        std::printf("\"zem_cols\": [");
        for (size_t k = 0; k < zemCols.size(); k++) std::printf(k ? ",%d" : "%d", zemCols[k]);
        std::printf("],\n\"sky_status\": %d,\n", sky.Status);
        for (size_t k = 0; k < zems.size(); k++) {
            std::string key = "zem_" + std::to_string(k);
            dump(key.c_str(), zems[k]);
        }
        dump("u_sky", sky.VectorX);
        std::vector<double> ke0(ke.begin(), ke.begin() + 300);
        dump("ke0", ke0, true);
        std::printf("}\n");
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
