// The MD objects of a configuration file, built on the host: the object
// factory of the reference's Simulation::add_md_object
// (src/simulation.cpp:330-351) for the types Clami, Sphere, Cilia and
// ConstantBeatCilia, producing one bead model that mpcd_md_attach takes.
#pragma once

#include <array>
#include <string>
#include <vector>

#include "cfg.hpp"
#include "mpcd_abi.h"

namespace mcm {

struct Vec3 {
    double x = 0.0, y = 0.0, z = 0.0;
};

// Bead-shell geometry of Sphere::create (src/sphere.cpp:137-245): the
// centre bead and a refined icosahedron (fill_factor 2: 162 surface beads,
// 3: 642) projected onto `radius` around `c`.  neilist[i][0] is the number
// of neighbours of surface bead i, neilist[i][1..] their indices (ascending);
// headbond the mean neighbour distance.
struct Shell {
    std::vector<Vec3> pts;
    std::vector<std::array<int, 7>> neilist;
    double headbond = 0.0;
};
Shell build_shell(int fill_factor, double radius, Vec3 c);

// Bead positions of Cilia::create (src/cilia.cpp:275-325): three rods of
// `rodlength` beads, rod-major.
std::vector<Vec3> build_cilium(int rodlength, Vec3 w, double thetaX, double thetaY, double thetaZ);

// What the run loop needs of each object, in the order the reference adds
// them to Simulation::md_objects (a Clami's sphere and cilia before it).
struct ObjectInfo {
    std::string type, name;
    bool write_data = true;     // MDObject::write_data (src/libcilia/mdobject.h:17-24)
    int first = 0, nbeads = 0;  // the object's beads in the bead model
    int sphere = -1;            // Sphere, Clami: index into BeadModel::sph
    std::vector<int> cilia;     // Cilia, Clami: indices into BeadModel::cil
    int att_first = 0, att_count = 0;  // Clami: its attachment springs
};

struct BeadModel {
    std::vector<double> pos3, vel4;
    std::vector<mpcd_md_sphere> sph;
    std::vector<int> nei_start, nei;
    std::vector<mpcd_md_cilium> cil;
    std::vector<double> newlength, oldlength;
    std::vector<mpcd_md_attachment> att;
    double att_k = 20000.0;
    bool att_k_set = false;
    std::vector<ObjectInfo> objects;

    int nbeads() const { return (int)(pos3.size() / 3); }
    // The descriptor over this model's arrays (valid while it lives).
    mpcd_md_desc desc(double globalt, double ststep) const;
};

// Simulation::add_md_object: adds the object a `md_objects` entry
// describes.  Throws CfgError on an unknown type or a bad setting.
void add_md_object(BeadModel &m, const Setting &obj);

}  // namespace mcm
