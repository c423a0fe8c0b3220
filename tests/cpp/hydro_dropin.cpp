// Drives the C++ drop-in (include/bemuse_hydro.h) the way the reference
// console drives Hydrodynamic_Radiation_Solver (BEMUser_Console.cpp:227-260):
// boundary parameters + Setup, solver parameters + Setup (which includes the
// FreqInf solve), then Set_Real(F) + Solve() per frequency. Prints one line per
// result for the test to compare with the oracle:
//   SETUP A <36>        A after Setup (Frequency = FreqInf)
//   A <omega> <36>      B <omega> <36>      F <omega> <12 nb>
// Usage: hydro_dropin <mode: 0 FP64 | 1 mixed> [ifr: 0 | 1]
// With ifr = 1 the cylinder gets a 3-ring interior lid
// (NPanels_FreeSurface_Int_Radial) and the solver the "IFR" key, and the
// frequencies are 1.0 and 2.2 (the irregular frequency of this cylinder).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bemuse_hydro.h"

using namespace BEMUse;

static void print_real(const char* tag, double w, const Matrix& M, bool with_w) {
    if (with_w) std::printf("%s %.17g", tag, w); else std::printf("%s", tag);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) std::printf(" %.17g", M(i, j));
    std::printf("\n");
}

int main(int argc, char** argv) {
    int mode = argc > 1 ? std::atoi(argv[1]) : BEM_MODE_FP64;
    bool ifr = argc > 2 && std::atoi(argv[2]) != 0;
    Half_Cylinder* B = new Half_Cylinder;
    std::vector<Parameter> bp = {Parameter("Radius", 5.0), Parameter("Draft", 8.0), Parameter("NPanels_Radial", 4),
                                 Parameter("NPanels_Vertical", 8), Parameter("NPanels_Azimuthal", 16),
                                 Parameter("Cosine_Disc", true)};
    if (ifr) bp.push_back(Parameter("NPanels_FreeSurface_Int_Radial", 3));
    B->Set_Parameters(bp);
    B->Setup();
    Hydrodynamic_Radiation_Solver* S = new Hydrodynamic_Radiation_Solver(0, mode);
    std::vector<Parameter> sp = {Parameter("WaveAngle", 0.0), Parameter("WaveAngle", 0.7853981633974483)};
    if (ifr) sp.push_back(Parameter("IFR", true));
    S->Set_Parameters(sp);
    std::string out = "dropin_test";
    S->Set_OutputFilePath(out);
    S->Setup(B);
    std::printf("N %d\n", S->N_Panels());
    print_real("SETUP_A", 0, S->Added_Mass(), false);
    std::vector<double> ws = ifr ? std::vector<double>{1.0, 2.2} : std::vector<double>{0.4, 1.3};
    for (double w : ws) {
        S->Set_Real(w * 2.0 * 3.14159265358979323846);   // the console passes F = 2 pi omega
        S->Solve();
        print_real("A", w, S->Added_Mass(), true);
        print_real("B", w, S->Damping(), true);
        std::printf("F %.17g", w);
        const CMatrix& F = S->Excitation_Force();
        for (int i = 0; i < 6; i++)
            for (int b = 0; b < F.cols(); b++) std::printf(" %.17g %.17g", F(i, b).real(), F(i, b).imag());
        std::printf("\n");
    }
    delete S;
    delete B;
    return 0;
}
