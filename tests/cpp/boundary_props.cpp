// Host-only checks of the drop-in's boundary properties and output field
// format (no GPU): prints name value pairs for tests/test_abi.py.
#include <cstdio>
#include <vector>

#include "bemuse_hydro.h"

using namespace BEMUse;

static void dump(const char* tag, Boundary& B) {
    std::printf("%s_volume %.17g\n%s_area %.17g\n%s_cob_x %.17g\n%s_cob_y %.17g\n%s_cob_z %.17g\n", tag, B.Volume, tag,
                B.Surface_Area, tag, B.Centre_Buoyancy.x, tag, B.Centre_Buoyancy.y, tag, B.Centre_Buoyancy.z);
    std::printf("%s_wp_area %.17g\n%s_wp_mx %.17g\n%s_wp_my %.17g\n%s_wp_mxx %.17g\n%s_wp_mxy %.17g\n%s_wp_myy %.17g\n",
                tag, B.WP_Area, tag, B.WP_Mx, tag, B.WP_My, tag, B.WP_Mxx, tag, B.WP_Mxy, tag, B.WP_Myy);
}

int main() {
    Barge box;
    std::vector<Parameter> bp = {Parameter("Tank_Length", 20.0), Parameter("Tank_Width", 10.0),
                                 Parameter("Tank_Depth", 5.0), Parameter("NPanels_Tank_Length", 10),
                                 Parameter("NPanels_Tank_Width", 5), Parameter("NPanels_Tank_Depth", 3)};
    box.Set_Parameters(bp);
    box.Setup();
    dump("box", box);
    Half_Cylinder cyl;
    std::vector<Parameter> cp = {Parameter("Radius", 5.0), Parameter("Draft", 8.0), Parameter("NPanels_Radial", 4),
                                 Parameter("NPanels_Vertical", 8), Parameter("NPanels_Azimuthal", 16)};
    cyl.Set_Parameters(cp);
    cyl.Setup();
    dump("cyl", cyl);
    std::printf("box_aux %d\ncyl_aux %d\n", box.N_Aux_Panels(), cyl.N_Aux_Panels());
    Half_Cylinder lidded;
    cp.push_back(Parameter("NPanels_FreeSurface_Int_Radial", 3));
    lidded.Set_Parameters(cp);
    lidded.Setup();
    std::printf("lidded_aux %d\n", lidded.N_Aux_Panels());
    std::printf("field_pos [%s]\nfield_neg [%s]\nfield_small [%s]\nfield_txt [%s]\n", Field(1.5).c_str(),
                Field(-2.25).c_str(), Field(1e-7).c_str(), Field(std::string("Surge: "), 30).c_str());
    return 0;
}
