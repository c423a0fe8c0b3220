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


// Mirror checks of Get_Symm_Mesh / Get_Aux_Symm_Mesh: node map exact (z
// negated, element order reversed), and the image, rebuilt as a mesh, encloses
// the same volume with its normals still pointing into it (positive volume).
static void mirror(const char* tag, const std::vector<Real>& n0, const std::vector<int>& e0,
                   const std::vector<Real>& n1, const std::vector<int>& e1) {
    int bad = (n0.size() != n1.size() || e0.size() != e1.size()) ? 1 : 0;
    for (size_t i = 0; !bad && i < n0.size(); i++) bad |= (i % 3 == 2 ? n1[i] != -n0[i] : n1[i] != n0[i]);
    for (size_t e = 0; !bad && e + 3 < e0.size(); e += 4) {
        if (e0[e + 3] < 0) bad |= !(e1[e] == e0[e + 2] && e1[e + 1] == e0[e + 1] && e1[e + 2] == e0[e] && e1[e + 3] < 0);
        else bad |= !(e1[e] == e0[e + 3] && e1[e + 1] == e0[e + 2] && e1[e + 2] == e0[e + 1] && e1[e + 3] == e0[e]);
    }
    double vol = 0, vol0 = 0;
    for (int k = 0; k < 2; k++) {
        const std::vector<Real>& nn = k ? n1 : n0;
        const std::vector<int>& ee = k ? e1 : e0;
        bem_mesh* m = nullptr;
        int ne = (int)(ee.size() / 4);
        if (ne == 0 || bem_mesh_from_arrays((int)(nn.size() / 3), nn.data(), ne, ee.data(), &m) != BEM_OK) continue;
        std::vector<double> body(32 * (size_t)ne), gen(8 * (size_t)ne);
        if (bem_geometry(m, body.data(), nullptr, nullptr, nullptr, gen.data()) == BEM_OK) {
            double v = 0;
            for (int i = 0; i < ne; i++) {
                const double* c = &body[32 * (size_t)i];
                const double* g = &gen[8 * (size_t)i];
                v -= (c[0] * g[0] + c[1] * g[1] + c[2] * g[2]) * g[6] / 3.0;   // normals point into the body
            }
            (k ? vol : vol0) = v;
        }
        bem_mesh_free(m);
    }
    std::printf("%s_mirror_bad %d\n%s_mirror_vol %.17g\n%s_vol %.17g\n", tag, bad, tag, vol, tag, vol0);
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
    {
        std::vector<Real> n0, n1;
        std::vector<int> e0, e1;
        lidded.Get_Mesh(n0, e0);
        lidded.Get_Symm_Mesh(n1, e1);
        mirror("hull", n0, e0, n1, e1);
        lidded.Get_Aux_Mesh(n0, e0);
        lidded.Get_Aux_Symm_Mesh(n1, e1);
        mirror("lid", n0, e0, n1, e1);
        Hydrodynamic_Radiation_Solver S;
        std::printf("vismat_rows %d\nvismat_cols %d\n", (int)S.Get_VisMatrix().rows(), (int)S.Get_VisMatrix().cols());
    }
    // reference parametric bodies (discretisations kept small)
    OC3_SparBuoy oc3;
    std::vector<Parameter> o3 = {Parameter("NPanels_Radial", 4), Parameter("NPanels_Vertical_Section1", 12),
                                 Parameter("NPanels_Vertical_Section2", 3), Parameter("NPanels_Vertical_Section3", 4),
                                 Parameter("NPanels_Azimuthal", 16), Parameter("NPanels_FreeSurface_Int_Radial", 2)};
    oc3.Set_Parameters(o3);
    oc3.Setup();
    dump("oc3", oc3);
    std::printf("oc3_aux %d\noc3_mass %.17g\noc3_m33 %.17g\noc3_c33 %.17g\n", oc3.N_Aux_Panels(), oc3.Mass,
                oc3.Mass_Mat(3, 3), oc3.Stiffness_Mat(2, 2));
    SemiSub_OC4 oc4;
    std::vector<Parameter> o4 = {Parameter("NPanels_Azimuthal", 24), Parameter("NPanels_Radial", 3),
                                 Parameter("NPanels_Vertical", 6), Parameter("NPanels_Radial_Base", 3),
                                 Parameter("NPanels_Vertical_HeavePlate", 2), Parameter("NPanels_Horizontal_HeavePlate", 2),
                                 Parameter("NPanels_FreeSurface_Int_Radial", 2), Parameter("NPanels_FreeSurface_Radial", 2),
                                 Parameter("FreeSurface_Radius", 1.0), Parameter("Scaling_Factor", 2.0)};
    oc4.Set_Parameters(o4);
    oc4.Setup();
    dump("oc4", oc4);
    std::printf("oc4_n %d\noc4_aux %d\noc4_fs %d\noc4_mass %.17g\noc4_c33 %.17g\n", oc4.N_Elements(),
                oc4.N_Aux_Panels(), (int)oc4.FreeSurface_Nodes.size(), oc4.Mass, oc4.Stiffness_Mat(2, 2));
    Triple_Spar ts;
    std::vector<Parameter> tp = {Parameter("Radius", 2.0), Parameter("Draft", 6.0), Parameter("NPanels_Radial", 2),
                                 Parameter("NPanels_Vertical", 3), Parameter("NPanels_Azimuthal", 12),
                                 Parameter("Spar_Radius", 10.0), Parameter("Shift_Over_Leg", true)};
    ts.Set_Parameters(tp);
    ts.Setup();
    dump("ts", ts);
    Semi_Ellipsoid se;
    std::vector<Parameter> sp = {Parameter("Semiaxis_a", 3.0), Parameter("Semiaxis_b", 2.0), Parameter("Semiaxis_c", 1.5),
                                 Parameter("NPanels_Axial", 6), Parameter("NPanels_Azimuthal", 12),
                                 Parameter("NPanels_FreeSurface_Int_Radial", 2), Parameter("NPanels_FreeSurface_Radial", 3)};
    se.Set_Parameters(sp);
    se.Setup();
    dump("se", se);
    std::printf("se_aux %d\nse_fs %d\n", se.N_Aux_Panels(), (int)se.FreeSurface_Nodes.size());
    Ellipsoid el;
    std::vector<Parameter> ep = {Parameter("Semiaxis_a", 3.0), Parameter("Semiaxis_b", 2.0), Parameter("Semiaxis_c", 1.5),
                                 Parameter("Ellipsoid_Depth", 5.0), Parameter("NPanels_Axial", 6),
                                 Parameter("NPanels_Azimuthal", 12)};
    el.Set_Parameters(ep);
    el.Setup();
    dump("el", el);
    Wigley_Hull wh;
    std::vector<Parameter> wp = {Parameter("Length", 10.0), Parameter("Beam", 1.0), Parameter("Draft", 0.625),
                                 Parameter("NPanels_Length", 20), Parameter("NPanels_Depth", 8)};
    wh.Set_Parameters(wp);
    wh.Setup();
    dump("wh", wh);
    std::printf("wh_n %d\nwh_fs %d\nwh_aux %d\n", wh.N_Elements(), (int)wh.FreeSurface_Nodes.size(), wh.N_Aux_Panels());
    Barge bf;
    std::vector<Parameter> bfp = {Parameter("Tank_Length", 20.0), Parameter("Tank_Width", 10.0),
                                  Parameter("Tank_Depth", 5.0), Parameter("NPanels_Tank_Length", 10),
                                  Parameter("NPanels_Tank_Width", 5), Parameter("NPanels_Tank_Depth", 3),
                                  Parameter("FreeSurface_Length", 40.0), Parameter("FreeSurface_Width", 20.0),
                                  Parameter("NPanels_FreeSurface_Length", 8), Parameter("NPanels_FreeSurface_Width", 4)};
    bf.Set_Parameters(bfp);
    bf.Setup();
    std::printf("bf_fs %d\nbf_fs_x0 %.17g\n", (int)bf.FreeSurface_Nodes.size(), bf.FreeSurface_Nodes[0].x);
    std::printf("field_pos [%s]\nfield_neg [%s]\nfield_small [%s]\nfield_txt [%s]\n", Field(1.5).c_str(),
                Field(-2.25).c_str(), Field(1e-7).c_str(), Field(std::string("Surge: "), 30).c_str());
    return 0;
}
