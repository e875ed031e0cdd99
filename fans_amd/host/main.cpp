// Drop-in CLI:  FANS <input.json> <output.h5>   (single rank, one MI355X)
//
// Same arguments, exit codes and step loop as the reference's src/main.cpp
// (runSolver :10-41, main :43-79): for every load case a fresh
// MaterialManager + Solver, for every time step set the gradient (or the mixed
// boundary condition), solve, postprocess, extrapolate the displacement.
// Environment: FANS_PRECISION=mixed selects the mixed FP32/FP64 form for
// linear problems (default: FP64, the drop-in form).
#include "reader.h"
#include "solver.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>

using namespace fans_host;

template <int howmany, int n_str>
static void runSolver(Reader &reader, int precision)
{
    for (size_t li = 0; li < reader.load_cases.size(); ++li) {
        const LoadCase                 &lc = reader.load_cases[li];
        MaterialManager<howmany, n_str> matmanager(reader);
        auto                            solver_ptr = createSolver<howmany, n_str>(reader, matmanager, precision);
        Solver<howmany, n_str>         &solver     = *solver_ptr;
        std::printf("\n# Load case %zu/%zu: %d time steps\n", li + 1, reader.load_cases.size(), lc.n_steps);
        for (int t = 0; t < lc.n_steps; ++t) {
            std::printf("#   Time step %d/%d (load case %zu/%zu)\n", t + 1, lc.n_steps, li + 1,
                        reader.load_cases.size());
            if (lc.mixed)
                solver.enableMixedBC(lc.mbc, (size_t) t);
            else
                matmanager.set_gradient(lc.g0_path[t]);
            solver.solve();
            solver.postprocess(reader, (int) li, t);
            if (reader.extrapolate_displacement)
                solver.extrapolateDisplacement();
        }
    }
}

int main(int argc, char *argv[])
{
    std::setvbuf(stdout, nullptr, _IOLBF, 0);
    if (argc > 1 && std::string(argv[1]) == "--version") {
        std::printf("FANS (MI355X) C ABI %d\n", fans_gpu_abi_version());
        return 0;
    }
    if (argc != 3) {
        std::fprintf(stderr, "USAGE: %s [input file basename] [output file basename]\n", argv[0]);
        return 10;
    }
    Reader reader;
    try {
        reader.ReadInputFile(argv[1]);
        reader.ReadMS();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "ERROR trying to read input file '%s' for FANS: %s\n", argv[1], e.what());
        return 10;
    }
    std::printf("# microstructure file name: \t '%s'\n", reader.ms_filename.c_str());
    std::printf("# microstructure dataset name: \t '%s'\n", reader.ms_datasetname.c_str());
    std::printf("# problem type: \t %s\n# FE type: \t %s\n", reader.problemType.c_str(), reader.FE_type.c_str());
    std::printf("# FANS error measure: \t %s %s error\n# FANS Tolerance: \t %10.5e\n# Max iterations: \t %6i\n",
                reader.err_type.c_str(), reader.err_measure.c_str(), reader.TOL, reader.n_it);
    std::printf("# grid size set to [%i x %i x %i] --> %zu voxels\n", reader.dims[0], reader.dims[1], reader.dims[2],
                reader.ms.size());
    const char *pe        = std::getenv("FANS_PRECISION");
    const int   precision = (pe && std::strcmp(pe, "mixed") == 0) ? FANS_PREC_MIXED : FANS_PREC_FP64;
    try {
        reader.OpenResultsFile(argv[2]);
        if (reader.howmany == 1)
            runSolver<1, 3>(reader, precision);
        else if (reader.n_str == 9)
            runSolver<3, 9>(reader, precision);
        else
            runSolver<3, 6>(reader, precision);
        reader.CloseResultsFile();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "ERROR: %s\n", e.what());
        return 1;
    }
    return 0;
}
