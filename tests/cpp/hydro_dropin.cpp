// Drives the C++ drop-in (include/bemuse_hydro.h) the way the reference
// console drives Hydrodynamic_Radiation_Solver (BEMUser_Console.cpp:227-260):
// boundary parameters + Setup, solver parameters + Setup (which includes the
// FreqInf solve), then Set_Real(F) + Solve() per frequency. Prints one line per
// result for the test to compare with the oracle:
//   SETUP A <36>        A after Setup (Frequency = FreqInf)
//   A <omega> <36>      B <omega> <36>      F <omega> <12 nb>
// Usage: hydro_dropin <mode: 0 FP64 | 1 mixed> [ifr: 0 | 1] [body: cyl | oc4]
// The plain cylinder (ifr = 0) has five exterior wave-elevation points.
// With ifr = 1 the body gets interior lids (NPanels_FreeSurface_Int_Radial)
// and the solver the "IFR" key; the cylinder's frequencies are then 1.0 and
// 2.2 (its irregular frequency). body oc4 is the OC4 semi-submersible
// (SemiSub_OC4) at a coarse discretisation, solved at 0.5 and 1.2.
#include <cstdio>
#include <cstdlib>
#include <string>
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
    bool oc4 = argc > 3 && std::string(argv[3]) == "oc4";
    Boundary* B = oc4 ? static_cast<Boundary*>(new SemiSub_OC4) : static_cast<Boundary*>(new Half_Cylinder);
    std::vector<Parameter> bp = {Parameter("Radius", 5.0), Parameter("Draft", 8.0), Parameter("NPanels_Radial", 4),
                                 Parameter("NPanels_Vertical", 8), Parameter("NPanels_Azimuthal", 16),
                                 Parameter("Cosine_Disc", true)};
    if (oc4)
        bp = {Parameter("NPanels_Azimuthal", 16), Parameter("NPanels_Radial", 2), Parameter("NPanels_Vertical", 4),
              Parameter("NPanels_Radial_Base", 2), Parameter("NPanels_Vertical_HeavePlate", 1),
              Parameter("NPanels_Horizontal_HeavePlate", 2)};
    if (ifr) bp.push_back(Parameter("NPanels_FreeSurface_Int_Radial", oc4 ? 2 : 3));
    B->Set_Parameters(bp);
    B->Setup();
    // the plain cylinder also gets exterior wave-elevation points (the
    // reference takes them from numerical-wave-tank boundaries): .6r and
    // _WaveElPos.bem are written for them
    if (!oc4 && !ifr)
        B->FreeSurface_Nodes = {Vector3(8.0, 0.0, 0.0), Vector3(0.0, 12.0, 0.0), Vector3(-15.0, 3.0, 0.0),
                                Vector3(10.0, -10.0, 0.0), Vector3(30.0, 20.0, 0.0)};
    Hydrodynamic_Radiation_Solver* S = new Hydrodynamic_Radiation_Solver(0, mode);
    std::vector<Parameter> sp = {Parameter("WaveAngle", 0.0), Parameter("WaveAngle", 0.7853981633974483)};
    if (ifr) sp.push_back(Parameter("IFR", true));
    S->Set_Parameters(sp);
    std::string out = "dropin_test";
    S->Set_OutputFilePath(out);
    S->Setup(B);
    std::printf("N %d\n", S->N_Panels());
    print_real("SETUP_A", 0, S->Added_Mass(), false);
    std::vector<double> ws = oc4 ? std::vector<double>{0.5, 1.2}
                                 : (ifr ? std::vector<double>{1.0, 2.2} : std::vector<double>{0.4, 1.3});
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
