// Object factory of the configuration driver (see objects.hpp).
#include "objects.hpp"

#include <cmath>

namespace mcm {

namespace {

constexpr double kPi = 3.14159265358979323846;  // PI, src/libcilia/abbrev.cpp:8
constexpr double kBondlength = 0.5;             // Cilia::bondlength, src/cilia.h:78
constexpr double kPartmass = 5.0;               // MDObject::partmass, src/libcilia/mdobject.h:76
constexpr int kAttachDepth = 3;                 // Clami::create_cilia `wo`, src/clami.cpp:124

double sqdist(const Vec3 &a, const Vec3 &b) {  // Matrix::sqdistance order (src/libcilia/Matrix.h:412-419)
    double s = 0.0, d;
    d = a.x - b.x;
    s += d * d;
    d = a.y - b.y;
    s += d * d;
    d = a.z - b.z;
    s += d * d;
    return s;
}

// Sphere::create_neillist (src/sphere.cpp:247-268): beads 1..count, pairs
// closer than 1.1.
void shell_neighbours(const std::vector<Vec3> &p, int count, std::vector<std::array<int, 7>> &nl) {
    const double rcutsq = 1.1 * 1.1;
    for (int j = 0; j <= count; ++j) nl[j].fill(0);
    for (int j = 1; j < count; ++j)
        for (int k = j + 1; k <= count; ++k)
            if (sqdist(p[j], p[k]) < rcutsq) {
                if (nl[j][0] == 6 || nl[k][0] == 6) throw CfgError("sphere mesh: a bead with more than 6 neighbours");
                nl[j][++nl[j][0]] = k;
                nl[k][++nl[k][0]] = j;
            }
}

const Setting &need(const Setting &g, const char *key) { return g[key]; }
double num(const Setting &g, const char *key) { return g[key].as_double(); }

Vec3 position_of(const Setting &g) {
    const Setting &p = need(g, "position");
    return Vec3{p[0].as_double(), p[1].as_double(), p[2].as_double()};
}

int append_beads(BeadModel &m, const std::vector<Vec3> &pts) {
    int first = m.nbeads();
    for (const Vec3 &q : pts) {
        m.pos3.insert(m.pos3.end(), {q.x, q.y, q.z});
        m.vel4.insert(m.vel4.end(), {0.0, 0.0, 0.0, kPartmass});  // MDObject::init (mdobject.cpp:152-159)
    }
    return first;
}

int add_sphere(BeadModel &m, const Shell &sh, double radius, double springconstant) {
    const int npart = (int)sh.pts.size();
    mpcd_md_sphere s{};
    s.first = append_beads(m, sh.pts);
    s.npart = npart;
    s.nei_off = (int)m.nei_start.size();
    s.radius = radius;
    s.springconstant = springconstant;
    s.headbond = sh.headbond;
    m.nei_start.push_back((int)m.nei.size());
    for (int i = 0; i < npart; ++i) {
        for (int j = 1; j <= sh.neilist[i][0]; ++j) m.nei.push_back(sh.neilist[i][j]);
        m.nei_start.push_back((int)m.nei.size());
    }
    m.sph.push_back(s);
    return (int)m.sph.size() - 1;
}

struct CiliumSpec {
    std::string type, name;
    Vec3 w;
    double thetaX, thetaY, thetaZ, start_delay;
    bool write_data;
};

// Cilia(cfg, ...) / ConstantBeatCilia(cfg, ...) (src/cilia.h:11-16,
// src/cilia.cpp:11-39, src/ConstantBeatCilia.h:14-21): startidx 0,
// num_anker 2, midpnt = num_anker, rest lengths = bondlength.
int add_cilium(BeadModel &m, const Setting &cfg, const CiliumSpec &c) {
    const long long rl = cfg["rodlength"].as_int();
    if (rl < 2 || rl > 4096) throw CfgError("setting " + cfg["rodlength"].path + " is out of range");
    const int R = (int)rl;
    const bool constant = c.type == "ConstantBeatCilia";
    mpcd_md_cilium k{};
    k.rodlength = R;
    k.num_anker = 2;
    k.constant_beat = constant ? 1 : 0;
    k.bondlength = kBondlength;
    k.beatstrength = num(cfg, "beatstrength");
    k.bending = num(cfg, "bending");
    k.min_bend = num(cfg, "min_bend");
    k.max_bend = num(cfg, "max_bend");
    k.m_speed = num(cfg, "m_speed");
    k.springconstant = num(cfg, "springconstant");
    if (constant) {
        k.power_stroke_duration = num(cfg, "power_stroke_duration");
        k.recovery_stroke_duration = num(cfg, "recovery_stroke_duration");
        k.Tstroke = c.start_delay;
    }
    k.midpnt = 2.0;
    k.first = append_beads(m, build_cilium(R, c.w, c.thetaX, c.thetaY, c.thetaZ));
    m.cil.push_back(k);
    m.newlength.insert(m.newlength.end(), R, kBondlength);
    m.oldlength.insert(m.oldlength.end(), R, kBondlength);
    ObjectInfo o;
    o.type = c.type;
    o.name = c.name;
    o.write_data = c.write_data;
    o.first = k.first;
    o.nbeads = 3 * R;
    o.cilia.push_back((int)m.cil.size() - 1);
    m.objects.push_back(o);
    return (int)m.cil.size() - 1;
}

bool cilium_type(const std::string &t) { return t == "Cilia" || t == "ConstantBeatCilia"; }

CiliumSpec cilium_from_cfg(const Setting &cfg, const std::string &type) {
    CiliumSpec c;
    c.type = type;
    c.name = cfg["name"].as_string();
    c.w = position_of(cfg);
    c.thetaX = num(cfg, "thetaX");
    c.thetaY = num(cfg, "thetaY");
    c.thetaZ = num(cfg, "thetaZ");
    c.start_delay = type == "ConstantBeatCilia" ? num(cfg, "start_delay") : 0.0;
    c.write_data = cfg["write_data"].as_bool();
    return c;
}

// Clami::add_springs_for_rods_idx (src/clami.cpp:136-160): rod bead `wo` of
// each rod to the closest shell bead (the first of equal distance,
// Matrix::get_closest_idx, src/libcilia/Matrix.h:602-614, over every bead
// of the shell including the centre) and to all of that bead's neighbours
// but the last, each spring at its present length.
void attach_rod_beads(BeadModel &m, const Shell &sh, int sph_first, int cil_first, int R, int wo) {
    for (int j = 0; j < 3; ++j) {
        const int b = cil_first + j * R + wo;
        const Vec3 q{m.pos3[3 * b], m.pos3[3 * b + 1], m.pos3[3 * b + 2]};
        double best = 9999.0;
        int idx = 0;
        for (int i = 0; i < (int)sh.pts.size(); ++i) {
            double d = sqdist(sh.pts[i], q);
            if (best > d) {
                best = d;
                idx = i;
            }
        }
        m.att.push_back(mpcd_md_attachment{sph_first + idx, b, std::sqrt(sqdist(sh.pts[idx], q))});
        for (int n = 1; n < sh.neilist[idx][0]; ++n) {
            const int s = sh.neilist[idx][n];
            m.att.push_back(mpcd_md_attachment{sph_first + s, b, std::sqrt(sqdist(sh.pts[s], q))});
        }
    }
}

// Clami::Clami and Clami::create_cilia (src/clami.cpp:32-134).
void add_clami(BeadModel &m, const Setting &cfg) {
    const std::string name = cfg["name"].as_string();
    double radius = 0.0;
    cfg.lookup("radius", radius);
    if (!(radius > 0.0)) throw CfgError("Clami " + name + ": radius must be positive");
    const Vec3 pos = position_of(cfg);
    double springconstant = 20000.0;
    cfg.lookup("springconstant", springconstant);
    if (m.att_k_set && m.att_k != springconstant)
        throw CfgError("Clami " + name + ": every Clami of one run must share its springconstant");
    m.att_k = springconstant;
    m.att_k_set = true;

    const Shell sh = build_shell(3, radius, pos);  // SphereNoWrite: 643 beads
    const int first = m.nbeads();
    const int si = add_sphere(m, sh, radius, springconstant);
    ObjectInfo so;
    so.type = "Sphere";
    so.name = name + "_sphere";
    so.write_data = false;  // SphereNoWrite::writedata writes nothing
    so.first = first;
    so.nbeads = (int)sh.pts.size();
    so.sphere = si;
    m.objects.push_back(so);

    const Setting &eq = need(cfg, "num_equators");
    const Setting &yt = need(cfg, "y_thetas");
    const Setting &off = need(cfg, "theta_offset");
    if (eq.length() != yt.length() || eq.length() != off.length())
        throw CfgError("Clami " + name + ": num_equators, y_thetas and theta_offset differ in length");
    long long num_cilia = 0;
    for (int i = 0; i < yt.length(); ++i) {
        long long n = eq[i].as_int();
        if (n < 0 || n > 100000) throw CfgError("setting " + eq[i].path + " is out of range");
        num_cilia += n;
    }
    const bool delays = cfg.exists("start_delays");
    if (delays && cfg["start_delays"].length() != num_cilia)
        throw CfgError("Clami " + name + ": start_delays does not hold one delay per cilium");

    const Setting &ccfg = need(cfg, "cilia");
    const std::string ctype = ccfg["type"].as_string();
    if (!cilium_type(ctype)) throw CfgError("Clami " + name + ": cilia of type " + ctype + " are not cilia");
    if (ccfg["rodlength"].as_int() <= kAttachDepth)
        throw CfgError("Clami " + name + ": rodlength must exceed 3, the rods attach to the shell at beads 0 and 3");
    for (const char *k : {"name", "thetaX", "thetaY", "start_delay"}) need(ccfg, k);  // assigned by create_cilia
    position_of(ccfg);
    const double thetaZ = num(ccfg, "thetaZ");
    const bool write_data = ccfg["write_data"].as_bool();

    const int att_first = (int)m.att.size();
    std::vector<int> cil;
    int cil_idx = 0;
    for (int i = 0; i < yt.length(); ++i) {
        double thetaY = yt[i].as_double();
        thetaY *= kPi / 180.;
        const double theta_offset = off[i].as_double();
        const int Nx = (int)eq[i].as_int();
        for (int j = 0; j < Nx; ++j) {
            CiliumSpec c;
            c.type = ctype;
            c.name = name + "_cilia" + std::to_string(cil_idx);
            c.thetaX = theta_offset + 2 * kPi * j / Nx;
            c.thetaY = thetaY;
            c.thetaZ = thetaZ;
            c.start_delay = delays ? cfg["start_delays"][cil_idx].as_double() : 0.0;
            c.write_data = write_data;
            const double r = radius - kAttachDepth * kBondlength;
            c.w = Vec3{r * std::sin(c.thetaY) + pos.x, -r * std::cos(c.thetaY) * std::sin(c.thetaX) + pos.y,
                       r * std::cos(c.thetaX) * std::cos(c.thetaY) + pos.z};
            const int ci = add_cilium(m, ccfg, c);
            const mpcd_md_cilium &k = m.cil[ci];
            attach_rod_beads(m, sh, first, k.first, k.rodlength, 0);
            attach_rod_beads(m, sh, first, k.first, k.rodlength, kAttachDepth);
            cil.push_back(ci);
            ++cil_idx;
        }
    }
    ObjectInfo o;
    o.type = "Clami";
    o.name = name;
    o.write_data = true;
    o.first = first;
    o.nbeads = m.nbeads() - first;
    o.sphere = si;
    o.cilia = cil;
    o.att_first = att_first;
    o.att_count = (int)m.att.size() - att_first;
    m.objects.push_back(o);
}

// Sphere(cfg, ...) (src/sphere.h:14-18, src/sphere.cpp:12-41).
void add_standalone_sphere(BeadModel &m, const Setting &cfg) {
    const std::string name = cfg["name"].as_string();
    const double radius = num(cfg, "radius");
    const double k = num(cfg, "springconstant");
    const Vec3 w = position_of(cfg);
    const long long npart = cfg["npart"].as_int();
    int fill;
    if (npart == 163) fill = 2;
    else if (npart == 643) fill = 3;
    else throw CfgError("Sphere " + name + ": Wrong nparts has to 163 or 643");
    if (!(radius > 0.0)) throw CfgError("Sphere " + name + ": radius must be positive");
    const Shell sh = build_shell(fill, radius, w);
    ObjectInfo o;
    o.type = "Sphere";
    o.name = name;
    o.first = m.nbeads();
    o.nbeads = (int)sh.pts.size();
    o.sphere = add_sphere(m, sh, radius, k);
    m.objects.push_back(o);
}

}  // namespace

Shell build_shell(int fill_factor, double radius, Vec3 c) {
    Shell sh;
    std::vector<Vec3> &p = sh.pts;
    const int maxn = 12 * (1 << (2 * fill_factor)) / 1 + 2;  // 10 * 4^f + 2 surface beads, with room
    p.assign(maxn + 1, Vec3{});
    sh.neilist.assign(maxn + 1, std::array<int, 7>{});
    // the icosahedron of src/sphere.cpp:145-180: poles at +-1.13, rings at
    // +-0.5023 of five vertices each; a = 2 pi/5, b = pi - 4 pi/5
    const double a = 2.0 * kPi / 5, b = kPi - 2.0 * 2.0 * kPi / 5;
    const double ring[12][3] = {
        {0., 0., 1.13},
        {std::cos(a), std::sin(a), 0.50230},
        {-std::cos(b), std::sin(b), 0.50230},
        {-std::cos(b), -std::sin(b), 0.5023},
        {std::cos(a), -std::sin(a), 0.50230},
        {1.0, 0.0, 0.5023},
        {std::cos(b), std::sin(b), -0.5023},
        {-std::cos(a), std::sin(a), -0.5023},
        {-1.0, 0.0, -0.5023},
        {-std::cos(a), -std::sin(a), -0.5023},
        {std::cos(kPi - 4.0 * kPi / 5), -std::sin(b), -0.5023},
        {0., 0., -1.13},
    };
    const double unit = 0.8513;  // bonds near unity
    for (int i = 0; i < 12; ++i) p[i + 1] = Vec3{ring[i][0] * unit, ring[i][1] * unit, ring[i][2] * unit};
    int npart = 12, tel = 12;
    for (int level = 0; level < fill_factor; ++level) {
        shell_neighbours(p, npart, sh.neilist);
        // a midpoint per mesh edge (j < k), in order of j, then k
        for (int j = 1; j < npart; ++j)
            for (int k = j + 1; k < npart + 1; ++k)
                for (int l = 1; l < sh.neilist[j][0] + 1; ++l)
                    if (sh.neilist[j][l] == k) {
                        ++tel;
                        if (tel > maxn) throw CfgError("sphere mesh: too many beads");
                        p[tel] = Vec3{(p[j].x + p[k].x) / 2.0, (p[j].y + p[k].y) / 2.0, (p[j].z + p[k].z) / 2.0};
                    }
        npart = tel;
        for (int i = 1; i < npart + 1; ++i) p[i] = Vec3{p[i].x * 2.0, p[i].y * 2.0, p[i].z * 2.0};
    }
    shell_neighbours(p, npart, sh.neilist);
    ++npart;  // with the centre bead
    p.resize(npart);
    sh.neilist.resize(npart);
    for (int i = 1; i < npart; ++i) {  // a quarter turn about y: the first pole to +x
        const double xi = p[i].x, zi = p[i].z;
        p[i].x = zi;
        p[i].z = -xi;
    }
    for (int i = 1; i < npart; ++i) {  // onto the sphere, then to its place
        const double s = radius / std::sqrt(sqdist(p[i], p[0]));
        p[i] = Vec3{p[i].x * s + c.x, p[i].y * s + c.y, p[i].z * s + c.z};
    }
    p[0] = Vec3{p[0].x + c.x, p[0].y + c.y, p[0].z + c.z};
    double sum = 0.0, cnt = 0.0;
    for (int i = 1; i < npart; ++i)
        for (int j = 1; j < sh.neilist[i][0] + 1; ++j) {
            sum += std::sqrt(sqdist(p[i], p[sh.neilist[i][j]]));
            cnt++;
        }
    sh.headbond = sum / cnt;
    return sh;
}

std::vector<Vec3> build_cilium(int R, Vec3 w, double tx, double ty, double tz) {
    const double bl = kBondlength;
    const double l0 = bl * std::sin(kPi / 3.0) * 2. / 3.;  // centroid offset of the triangle
    std::vector<Vec3> p(3 * (size_t)R);
    for (int i = 0; i < R; ++i) {
        p[i] = Vec3{0.0 - l0, 0.0, bl * i};
        p[R + i] = Vec3{bl * std::sin(kPi / 3.0) - l0, bl / 2.0, bl * i};
        p[2 * R + i] = Vec3{bl * std::sin(kPi / 3.0) - l0, -bl / 2.0, bl * i};
    }
    const double sx = std::sin(tx), cx = std::cos(tx), sy = std::sin(ty), cy = std::cos(ty), sz = std::sin(tz),
                 cz = std::cos(tz);
    for (Vec3 &q : p) {  // the rotation of src/cilia.cpp:300-316, then the translation
        const double xi = q.x, yi = q.y, zi = q.z;
        const double inner = zi * cy + sy * (-(xi * cz) + yi * sz);
        const double yz = yi * cz + xi * sz;
        q.x = zi * sy + cy * (xi * cz - yi * sz) + w.x;
        q.y = cx * yz - sx * inner + w.y;
        q.z = sx * yz + cx * inner + w.z;
    }
    return p;
}

mpcd_md_desc BeadModel::desc(double globalt, double ststep) const {
    mpcd_md_desc d{};
    d.nb = nbeads();
    d.pos3 = pos3.data();
    d.vel4 = vel4.data();
    d.mobile = nullptr;  // startidx 0 for every object built from a configuration
    d.nsph = (int)sph.size();
    d.sph = sph.data();
    d.nei_start = nei_start.data();
    d.nei = nei.data();
    d.ncil = (int)cil.size();
    d.cil = cil.data();
    d.newlength = newlength.data();
    d.oldlength = oldlength.data();
    d.natt = (int)att.size();
    d.att = att.data();
    d.att_springconstant = att_k;
    d.globalt = globalt;
    d.ststep = ststep;
    return d;
}

void add_md_object(BeadModel &m, const Setting &obj) {
    const std::string type = obj["type"].as_string();
    if (type == "Clami") {
        add_clami(m, obj);
    } else if (type == "Sphere") {
        add_standalone_sphere(m, obj);
    } else if (cilium_type(type)) {
        add_cilium(m, obj, cilium_from_cfg(obj, type));
    } else {
        throw CfgError("Unkown type:" + type);
    }
}

}  // namespace mcm
