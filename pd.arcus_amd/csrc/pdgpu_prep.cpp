// pdgpu_prep.cpp -- input preparation of the path, on the host: PD's
// parameter-file reader (FFParamSet::readLib, forcefields/ffparam.cpp:
// 1719-2200), its sequence grammar (Sequence::BioSequence::append,
// sequence/sequence.cpp:165-452), the extended peptide builder NewProtein
// (system/genpolymer.cpp:12-164,256-270), the covalent 1-2 / 1-3 / 1-4
// structure (system/molecule.cpp:123-322), the phi / psi / omega setters
// (system/molecule.cpp:613-721,1099-1180) and the FF_Bonded term lists with
// their parameter lookup (forcefields/ffbonded.cpp:13-460,
// forcefields/ffparam.cpp:74-500,1479-1645).
//
// A PD user reads their own parameter file with it; none ships here. The
// outputs are what the device path consumes: per-atom arrays in PD units,
// the 1-2 / 1-3 / 1-4 pair lists with equilibrium distances, positions and
// the FF_Bonded terms in PD's order.
#include "../../include/pdgpu.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// PD's constants (forcefields/forcefield.cpp:18-38, maths/maths.cpp:6-10)
constexpr double kKcal2J = 4184.0;
constexpr double kNa = 6.0221418E23;
constexpr double kAmu = 1.66053878E-27;
constexpr double kPI = 3.14159265358979323846;
const double kPIOver180 = kPI / 180.0;

struct Vec {
    double x = 0, y = 0, z = 0;
};
inline Vec sub(const Vec &a, const Vec &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec cross(const Vec &u, const Vec &v) {
    return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
inline double dot(const Vec &u, const Vec &v) { return u.x * v.x + u.y * v.y + u.z * v.z; }
inline double mag(const Vec &u) { return std::sqrt(u.x * u.x + u.y * u.y + u.z * u.z); }

// Tvector::angleWith (maths/maths_vector.h:341-349)
double angle_with(const Vec &u, const Vec &v) {
    const double s = dot(u, v) / (mag(u) * mag(v));
    if (s > 1) return 0;
    if (s < -1) return kPI;
    return std::acos(s);
}

struct Mat {
    double r[3][3];
};
// matrix3x3::setToAxisRot (maths/maths_matrix3x3.cpp:161-185), same
// operation order
Mat axis_rot(const Vec &axis, double angle) {
    const double invmag = 1.0 / mag(axis);
    const double c = std::cos(angle), s = std::sin(angle), t = 1.0 - c;
    const double tx = t * axis.x * invmag;
    const double sx = s * axis.x * invmag, sy = s * axis.y * invmag, sz = s * axis.z * invmag;
    const double tyz = t * axis.y * invmag * axis.z * invmag;
    const double ay = axis.y * invmag, az = axis.z * invmag;
    Mat m;
    m.r[0][0] = tx * axis.x * invmag + c;
    m.r[0][1] = tx * axis.y * invmag - sz;
    m.r[0][2] = tx * axis.z * invmag + sy;
    m.r[1][0] = tx * axis.y * invmag + sz;
    m.r[1][1] = t * (ay * ay) + c;
    m.r[1][2] = tyz - sx;
    m.r[2][0] = tx * axis.z * invmag - sy;
    m.r[2][1] = tyz + sx;
    m.r[2][2] = t * (az * az) + c;
    return m;
}
// Tvector::mulmat (maths/maths_vector.h:371-380)
Vec mulmat(const Vec &v, const Mat &m) {
    return {m.r[0][0] * v.x + m.r[0][1] * v.y + m.r[0][2] * v.z, m.r[1][0] * v.x + m.r[1][1] * v.y + m.r[1][2] * v.z,
            m.r[2][0] * v.x + m.r[2][1] * v.y + m.r[2][2] * v.z};
}
// Maths::calcTorsionAngle (maths/maths.h:240-260)
double torsion_angle(const Vec &i, const Vec &a, const Vec &b, const Vec &j) {
    const Vec r12 = sub(i, a), r23 = sub(a, b), r34 = sub(b, j);
    const Vec A = cross(r12, r23), B = cross(r23, r34), C = cross(r23, A);
    const double rA = mag(A), rB = mag(B), rC = mag(C);
    const double cos_phi = dot(A, B) / (rA * rB);
    const double sin_phi = dot(C, B) / (rC * rB);
    return -std::atan2(sin_phi, cos_phi);
}

// --- the parameter set ----------------------------------------------------
struct AtomType {
    std::string name;
    int Z = 0;
    double mass = 0, radius = 0, epsilon = 0, charge = 0;
};
struct BondType {
    int i, j;
    double length, k;
};
struct AngleType {
    int i, a, j;
    double theta0, k;
};
struct TorsionType {
    int kind; // 0 torsion, 1 improper
    int i, a, b, j;
    int terms = 0;
    double Vn[4] = {0, 0, 0, 0}, n[4] = {0, 0, 0, 0}, gamma[4] = {0, 0, 0, 0};
};
struct Link {
    std::string name; // atom name (rawname) in residue ir + offset
    int offset = 0;
    int index = -1;   // within the molecule when offset == 0
};
struct TemplateAtom {
    std::string rawname, pdbname, type_name, restraint;
    int fftype = 0, Z = 0;
    Vec pos;
    double charge = 0, radius = 0, epsilon = 0, mass = 0;
    std::vector<Link> cov;
};
struct DihedralDef {
    std::string an[4];
    int ro[4] = {0, 0, 0, 0};
};
struct MolDef {
    char letter = '?';
    std::string l3, name;
    bool append_end = true; // ATOM_APPEND_MODE END (the default) or START
    bool backlink = false, frwdlink = false;
    std::string backname, frwdname;
    Vec backpos, frwdpos;
    std::vector<TemplateAtom> atom;
    std::vector<DihedralDef> improper;
    int find_raw(const std::string &s) const {
        for (size_t k = 0; k < atom.size(); ++k)
            if (atom[k].rawname == s) return (int)k;
        return -1;
    }
};
struct ResidueClass {
    std::string name; // lower case
    std::vector<std::string> members, caps;
    std::string start, end; // terminal codes
    bool start_post = false, end_post = false;
};

std::vector<std::string> chop(const std::string &s, const char *delim) {
    std::vector<std::string> t;
    size_t pos = 0;
    while (pos < s.size()) {
        pos = s.find_first_not_of(delim, pos);
        if (pos == std::string::npos) break;
        size_t e = s.find_first_of(delim, pos);
        if (e == std::string::npos) e = s.size();
        t.push_back(s.substr(pos, e - pos));
        pos = e;
    }
    return t;
}
const char *kSpace = " \t\n\r";

bool num(const std::string &s, double &v) { return std::sscanf(s.c_str(), "%lf", &v) == 1; }
bool inum(const std::string &s, int &v) { return std::sscanf(s.c_str(), "%d", &v) == 1; }
bool ident(const std::string &s) {
    for (char ch : s)
        if (!std::isalnum((unsigned char)ch) && ch != '_') return false;
    return true;
}
std::string lower(std::string s) {
    for (char &ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}
// interpretAtomReference (forcefields/ffparam.cpp:29-48)
void atom_ref(const std::string &id, std::string &name, int &offset) {
    if (!id.empty() && id[0] == '-') { name = id.substr(1); offset = -1; }
    else if (!id.empty() && id[0] == '+') { name = id.substr(1); offset = 1; }
    else { name = id; offset = 0; }
}

struct ParamSet {
    std::vector<AtomType> types;
    std::vector<BondType> bonds;
    std::vector<AngleType> angles;
    std::vector<TorsionType> torsions; // torsions and impropers, in file order
    std::vector<MolDef> mols;
    std::vector<std::pair<std::string, std::string>> aliases; // (alias, name)
    std::vector<ResidueClass> classes;
    std::map<std::string, std::vector<std::string>> sections;
    bool autobonds = true, autoangles = true, autotorsions = true;
    double vdw14 = 1.0, elec14 = 1.0;
    std::vector<double> gb_radius; // per atom type (SOLVATION_GENERALISED_BORN)
    double gb_p[5] = {-1, -1, -1, -1, -1};
    std::string error;

    // FFParamSet::findAtomType: -1 wildcard "X", -2 unknown
    int find_type(const std::string &s) const {
        if (s == "X") return -1;
        for (size_t k = 0; k < types.size(); ++k)
            if (types[k].name == s) return (int)k;
        return -2;
    }
    int find_mol_exact(const std::string &s) const {
        for (size_t k = 0; k < mols.size(); ++k)
            if (mols[k].name == s) return (int)k;
        return -1;
    }
    // FFParamSet::findMoleculeType: the name, else its alias
    int find_mol(const std::string &s) const {
        int k = find_mol_exact(s);
        if (k >= 0) return k;
        for (const auto &a : aliases)
            if (a.first == s) return find_mol_exact(a.second);
        return -1;
    }
    ResidueClass &klass(const std::string &name) {
        const std::string n = lower(name);
        for (auto &c : classes)
            if (c.name == n) return c;
        classes.push_back(ResidueClass());
        classes.back().name = n;
        return classes.back();
    }
    void read(const std::string &path, int level);
    void read_molecule(FILE *f, MolDef &m);
    void finish_gb();
};

// the wildcard-preferring lookups (forcefields/ffparam.cpp:111-175,
// 218-287, 289-365, 1479-1643): the candidate with the fewest wildcards, the
// first of equals; the reversed order matches too
int match_count(const int *pat, const int *t, int n) {
    for (int dir = 0; dir < 2; ++dir) {
        int result = n, wild = 0;
        for (int k = 0; k < n; ++k) {
            const int want = dir == 0 ? t[k] : t[n - 1 - k];
            if (pat[k] == -1) { --result; ++wild; }
            else if (pat[k] == want) --result;
        }
        if (result == 0) return wild;
    }
    return -1;
}
template <typename F> int best_match(size_t count, F &&score) {
    int cand = -1, wild = 4;
    for (size_t k = 0; k < count; ++k) {
        const int w = score(k);
        if (w < 0 || w == wild) continue;
        if (w < wild) { wild = w; cand = (int)k; }
    }
    return cand;
}

void ParamSet::read_molecule(FILE *f, MolDef &m) {
    char buf[65535];
    while (std::fgets(buf, sizeof(buf), f)) {
        std::string line = buf;
        const size_t c = line.find_first_of("#\n\r");
        if (c != std::string::npos) line = line.substr(0, c);
        const std::vector<std::string> t = chop(line, kSpace);
        if (t.empty()) continue;
        const std::string &cmd = t[0];
        if (cmd == "ENDMOLECULE" || cmd == "ENDAMINOACID") break;
        if (cmd == "LOAD") {
            if (t.size() < 2) continue;
            const int k = find_mol(t[1]);
            if (k < 0) continue;
            m = mols[k];
        } else if (cmd == "ATOM_APPEND_MODE") {
            if (t.size() < 2) continue;
            if (t[1] == "START") m.append_end = false;
            else if (t[1] == "END") m.append_end = true;
        } else if (cmd == "ATOM") {
            if (t.size() < 7) continue;
            TemplateAtom a;
            if (!ident(t[1]) || !ident(t[2])) continue;
            a.rawname = t[1];
            a.pdbname = t[2];
            if (!num(t[3], a.pos.x) || !num(t[4], a.pos.y) || !num(t[5], a.pos.z)) continue;
            a.type_name = t[6];
            if (t.size() >= 8) a.restraint = t[7];
            const int ty = find_type(a.type_name);
            if (ty < 0) continue;
            const AtomType &T = types[ty];
            a.fftype = ty;
            a.Z = T.Z;
            a.charge = T.charge;
            a.radius = T.radius;
            a.epsilon = T.epsilon;
            a.mass = T.mass;
            if (m.append_end) m.atom.push_back(a);
            else m.atom.insert(m.atom.begin(), a);
        } else if (cmd == "DELATOM") {
            if (t.size() < 2) continue;
            const int k = m.find_raw(t[1]);
            if (k >= 0) m.atom.erase(m.atom.begin() + k);
        } else if (cmd == "DELCONNECT") {
            if (t.size() < 2) continue;
            if (t[1][0] == '+') { m.frwdname.clear(); m.frwdlink = false; }
            else if (t[1][0] == '-') { m.backname.clear(); m.backlink = false; }
        } else if (cmd == "CONNECT") {
            if (t.size() < 5) continue;
            Vec p;
            if (!num(t[2], p.x) || !num(t[3], p.y) || !num(t[4], p.z)) continue;
            if (t[1][0] == '+') { m.frwdname = t[1].substr(1); m.frwdpos = p; }
            else if (t[1][0] == '-') { m.backname = t[1].substr(1); m.backpos = p; }
        } else if (cmd == "IMPROPER" || cmd == "DELIMPROPER") {
            if (t.size() < 5) continue;
            DihedralDef d;
            for (int k = 0; k < 4; ++k) atom_ref(t[1 + k], d.an[k], d.ro[k]);
            if (cmd == "IMPROPER") {
                m.improper.push_back(d);
            } else {
                for (size_t k = 0; k < m.improper.size(); ++k) {
                    const DihedralDef &q = m.improper[k];
                    if (q.an[0] == d.an[0] && q.an[1] == d.an[1] && q.an[2] == d.an[2] && q.an[3] == d.an[3]) {
                        m.improper.erase(m.improper.begin() + k);
                        break;
                    }
                }
            }
        } else if (cmd == "CHARGE" || cmd == "VDWRADIUS" || cmd == "EPSILON" || cmd == "MASS") {
            if (t.size() < 3) continue;
            const int k = m.find_raw(t[1]);
            double v;
            if (k < 0 || !num(t[2], v)) continue;
            if (cmd == "CHARGE") m.atom[k].charge = v;
            else if (cmd == "VDWRADIUS") m.atom[k].radius = v;
            else if (cmd == "EPSILON") { if (v >= 0) m.atom[k].epsilon = std::sqrt(v * kKcal2J / kNa); }
            else m.atom[k].mass = v * kAmu;
        }
        // TORSION lines inside molecules and custom properties are not
        // used by the path
    }
    // decodeRestraintList over every atom (forcefields/ffparam.cpp:548-583):
    // links to neighbouring residues set the link flags, unknown names are
    // skipped
    for (auto &a : m.atom) {
        a.cov.clear();
        for (const std::string &r : chop(a.restraint, ",")) {
            Link l;
            atom_ref(r, l.name, l.offset);
            if (l.offset != 0) {
                if (l.offset < 0) m.backlink = true;
                else m.frwdlink = true;
            } else {
                l.index = m.find_raw(l.name);
                if (l.index < 0) continue;
            }
            a.cov.push_back(l);
        }
    }
}

void ParamSet::read(const std::string &path, int level) {
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) throw std::runtime_error("parameter file '" + path + "' not found");
    std::unique_ptr<FILE, int (*)(FILE *)> guard(f, std::fclose);
    char buf[65535];
    while (std::fgets(buf, sizeof(buf), f)) {
        std::string line = buf;
        const size_t c = line.find_first_of("#\n\r");
        if (c != std::string::npos) line = line.substr(0, c);
        const std::vector<std::string> t = chop(line, kSpace);
        if (t.empty()) continue;
        const std::string &cmd = t[0];
        if (cmd == "INCLUDE" && t.size() >= 2) {
            std::string p = t[1];
            if (!(p.size() > 2 && p[1] == ':')) {
                const size_t s = path.find_last_of("\\/");
                p = (s == std::string::npos ? std::string() : path.substr(0, s + 1)) + t[1];
            }
            FILE *g = std::fopen(p.c_str(), "r");
            if (!g) continue; // PD warns and goes on
            std::fclose(g);
            read(p, level + 1);
        } else if (cmd == "ALIAS" && t.size() >= 3) {
            if (!ident(t[1]) || !ident(t[2]) || t[1] == t[2]) continue;
            for (int k = (int)aliases.size() - 1; k >= 0; --k)
                if (aliases[k].first == t[1]) aliases.erase(aliases.begin() + k);
            aliases.push_back({t[1], t[2]});
            // resolve chains of aliases (forcefields/ffparam.cpp readLib)
            for (bool changed = true; changed;) {
                changed = false;
                for (auto &a : aliases)
                    for (auto &b : aliases)
                        if (a.second == b.first) { a.second = b.second; changed = true; }
            }
        } else if (cmd == "CLASS" && t.size() >= 3) {
            klass(t[1]).members.push_back(t[2]);
        } else if (cmd == "CLASS_CAP" && t.size() >= 3) {
            klass(t[1]).caps.push_back(t[2]);
        } else if ((cmd == "CLASS_PREFIX" || cmd == "CLASS_POSTFIX") && t.size() >= 4) {
            ResidueClass &k = klass(t[1]);
            k.start = t[2];
            k.end = t[3];
            k.start_post = k.end_post = cmd == "CLASS_POSTFIX";
        } else if (cmd == "AUTOBONDS") autobonds = true;
        else if (cmd == "NOAUTOBONDS") autobonds = false;
        else if (cmd == "AUTOANGLES") autoangles = true;
        else if (cmd == "NOAUTOANGLES") autoangles = false;
        else if (cmd == "AUTOTORSIONS") autotorsions = true;
        else if (cmd == "NOAUTOTORSIONS") autotorsions = false;
        else if (cmd == "VDW14SCALING" && t.size() >= 2) num(t[1], vdw14);
        else if (cmd == "ELEC14SCALING" && t.size() >= 2) num(t[1], elec14);
        else if (cmd == "SECTION" && t.size() >= 2) {
            std::vector<std::string> &lines = sections[t[1]];
            while (std::fgets(buf, sizeof(buf), f)) {
                std::string s = buf;
                std::string u = s;
                const size_t cc = u.find_first_of("#\n\r");
                if (cc != std::string::npos) u = u.substr(0, cc);
                const std::vector<std::string> q = chop(u, kSpace);
                if (!q.empty() && q[0] == "ENDSECTION") break;
                lines.push_back(s);
            }
        } else if (cmd == "TYPE") {
            // AtomTypeParameter::AtomTypeReadLine (ffparam.cpp:1444-1477)
            if (t.size() < 7) { error = "TYPE line incomplete"; continue; }
            AtomType a;
            a.name = t[1];
            if (!inum(t[2], a.Z) || !num(t[3], a.mass) || !num(t[4], a.radius) || !num(t[5], a.epsilon) ||
                !num(t[6], a.charge) || a.epsilon < 0) {
                error = "TYPE line unreadable: " + line;
                continue;
            }
            a.epsilon = std::sqrt(a.epsilon * kKcal2J / kNa);
            a.mass *= kAmu;
            types.push_back(a);
        } else if (cmd == "BOND") {
            // BondTypeParameter::readDefinitionLine (ffparam.cpp:74-109)
            if (t.size() < 5) continue;
            BondType b;
            b.i = find_type(t[1]);
            b.j = find_type(t[2]);
            if (b.i < -1 || b.j < -1 || !num(t[3], b.length) || !num(t[4], b.k)) continue;
            b.k *= kKcal2J / kNa;
            bonds.push_back(b);
        } else if (cmd == "ANGLE") {
            if (t.size() < 6) continue;
            AngleType a;
            a.i = find_type(t[1]);
            a.a = find_type(t[2]);
            a.j = find_type(t[3]);
            if (a.i < -1 || a.a < -1 || a.j < -1 || !num(t[4], a.theta0) || !num(t[5], a.k)) continue;
            a.theta0 *= kPIOver180;
            a.k *= kKcal2J / kNa;
            angles.push_back(a);
        } else if (cmd == "TORSION" || cmd == "IMPROPER") {
            // TorsionTypeParameter::readDefinitionLine (ffparam.cpp:408-497):
            // a second line for the same quartet adds a Fourier term (or
            // replaces the one of equal periodicity)
            if (t.size() < 8) continue;
            TorsionType q;
            q.kind = cmd == "TORSION" ? 0 : 1;
            q.i = find_type(t[1]);
            q.a = find_type(t[2]);
            q.b = find_type(t[3]);
            q.j = find_type(t[4]);
            double vn, gamma, n;
            if (q.i < -1 || q.a < -1 || q.b < -1 || q.j < -1 || !num(t[5], vn) || !num(t[6], gamma) || !num(t[7], n))
                continue;
            vn *= kKcal2J / kNa;
            gamma *= kPIOver180;
            TorsionType *target = nullptr;
            for (auto &p : torsions)
                if (p.i == q.i && p.a == q.a && p.b == q.b && p.j == q.j && p.kind == q.kind) { target = &p; break; }
            const bool fresh = target == nullptr;
            if (fresh) target = &q;
            int slot = target->terms;
            for (int k = 0; k < target->terms; ++k)
                if (target->n[k] == n) { slot = k; break; }
            if (target->terms >= 4) throw std::runtime_error("more than 4 Fourier terms for one torsion");
            target->Vn[slot] = vn;
            target->n[slot] = n;
            target->gamma[slot] = gamma;
            if (slot == target->terms) target->terms++;
            if (fresh) torsions.push_back(q);
        } else if (cmd == "MOLECULE") {
            MolDef m;
            read_molecule(f, m);
            // the names are set after the block: LOAD overwrites them
            if (t.size() >= 4) {
                if (!ident(t[2]) || !ident(t[3])) continue;
                m.letter = t[1][0];
                m.l3 = t[2].substr(0, 3);
                m.name = t[3];
            } else if (t.size() >= 3) {
                if (!ident(t[2])) continue;
                m.letter = t[1][0];
                m.name = t[2];
                m.l3 = m.name.size() > 3 ? m.name.substr(m.name.size() - 3) : m.name;
            }
            mols.push_back(m);
        }
        // INFO, IDENT, CUSTOMPROPERTY and per-type custom properties carry
        // nothing the path reads
    }
    if (level == 0) finish_gb();
}

// FF_GeneralizedBorn_Still's parameter section (forcefields/gbff.cpp:
// 185-290): GENERALPARAMETERS P1..P5 and a Still radius per atom type
void ParamSet::finish_gb() {
    gb_radius.assign(types.size(), -1.0);
    auto it = sections.find("SOLVATION_GENERALISED_BORN");
    if (it == sections.end()) return;
    for (std::string line : it->second) {
        const size_t c = line.find_first_of("#\n\r");
        if (c != std::string::npos) line = line.substr(0, c);
        const std::vector<std::string> t = chop(line, kSpace);
        if (t.empty()) continue;
        if (t[0] == "GENERALPARAMETERS" && t.size() >= 6) {
            for (int k = 0; k < 5; ++k) num(t[1 + k], gb_p[k]);
        } else if (t[0] == "SOLVATIONTYPE" && t.size() >= 3) {
            const int ty = find_type(t[1]);
            if (ty >= 0) num(t[2], gb_radius[ty]);
        }
    }
}

// --- the built molecule ---------------------------------------------------
struct CovAtom {
    int i, i2 = -1, i3 = -1;
    double d = 0;
};
struct Particle {
    const TemplateAtom *tpl;
    int mol; // index of its residue's MolDef
    int ir;
    Vec pos, ref;
    std::vector<CovAtom> cov12, cov13, cov14;
};
struct Residue {
    int ifirst, ilast, iN = -1, iCA = -1, iC = -1, mol;
};

struct Molecule {
    const ParamSet *ps;
    std::vector<Particle> atom;
    std::vector<Residue> res;
    // FF_Bonded terms
    std::vector<int> bond_ij, angle_iaj, tors_iabj, imp_iabj, tors_terms, imp_terms;
    std::vector<double> bond_lk, angle_tk, tors_vn, tors_n, tors_gamma, imp_vn, imp_n, imp_gamma;

    int by_raw(int ir, const std::string &name) const {
        for (size_t k = 0; k < atom.size(); ++k)
            if (atom[k].ir == ir && atom[k].tpl->rawname == name) return (int)k;
        return -1;
    }
    int by_pdb(int ir, const std::string &name) const {
        for (size_t k = 0; k < atom.size(); ++k)
            if (atom[k].ir == ir && atom[k].tpl->pdbname == name) return (int)k;
        return -1;
    }
    void residues();
    void covalent();
    void calc1314();
    bool rotate_bond(int bi, int bj, double angle);
    void phipsi(int ir, double &phi, double &psi) const;
    void omega(int ir, double &om) const;
    void set_phipsi(int ir, double phi, double psi);
    void set_omega(int ir, double om);
    void bonded();
};

// MoleculeBase::detectResidueBoundaries (system/molecule.cpp:331-438): the
// last atom of a residue named N, CA or C by its PDB name is its N, CA, C
void Molecule::residues() {
    res.clear();
    int last = -1;
    for (size_t k = 0; k < atom.size(); ++k)
        if (atom[k].ir != last) {
            Residue r;
            r.ifirst = (int)k;
            r.ilast = (int)atom.size() - 1;
            r.mol = atom[k].mol;
            if (!res.empty()) res.back().ilast = (int)k - 1;
            res.push_back(r);
            last = atom[k].ir;
        }
    for (size_t k = 0; k < atom.size(); ++k) {
        const std::string &p = atom[k].tpl->pdbname;
        Residue &r = res[atom[k].ir];
        if (p == "CA") r.iCA = (int)k;
        if (p == "N") r.iN = (int)k;
        if (p == "C") r.iC = (int)k;
    }
}

// MoleculeBase::createCovalentStructure + calc1314BondOrders
// (system/molecule.cpp:123-157,237-322)
void Molecule::covalent() {
    for (size_t ir = 0; ir < res.size(); ++ir) {
        const MolDef &m = ps->mols[res[ir].mol];
        for (const TemplateAtom &ta : m.atom) {
            const int t = by_raw((int)ir, ta.rawname);
            for (const Link &l : ta.cov) {
                const int p = by_raw((int)ir + l.offset, l.name);
                if (p < 0)
                    throw std::runtime_error("covalent structure: atom " + l.name + " of residue " +
                                             std::to_string((int)ir + l.offset) + " not found");
                CovAtom c;
                c.i = p;
                atom[t].cov12.push_back(c);
            }
        }
    }
    // checkCovalentStructure: every bond recorded from both ends
    for (size_t j = 0; j < atom.size(); ++j)
        for (const CovAtom &c : atom[j].cov12) {
            bool ok = false;
            for (const CovAtom &b : atom[c.i].cov12) ok |= b.i == (int)j;
            if (!ok) throw std::runtime_error("covalent structure is not symmetric");
        }
    calc1314();
}

void Molecule::calc1314() {
    struct Trail {
        int i1, i2, i3, i4;
        int order;
    };
    for (Particle &p : atom) { p.cov13.clear(); p.cov14.clear(); }
    std::vector<Trail> list;
    for (size_t i = 0; i < atom.size(); ++i) {
        list.clear();
        list.push_back({(int)i, -1, -1, -1, 0});
        for (size_t li = 0; li < list.size(); ++li) {
            const Trail e = list[li];
            if (e.order == 2) {
                CovAtom c;
                c.i = e.i1;
                c.i2 = e.i2;
                atom[i].cov13.push_back(c);
            } else if (e.order == 3) {
                CovAtom c;
                c.i = e.i1;
                c.i2 = e.i2;
                c.i3 = e.i3;
                atom[i].cov14.push_back(c);
            }
            if (e.order >= 3) continue;
            for (const CovAtom &c : atom[e.i1].cov12) {
                if (c.i == e.i2) continue; // no step back
                list.push_back({c.i, e.i1, e.i2, e.i3, e.order + 1});
                if (list.size() >= 10 * 4 * 3 * 3) throw std::runtime_error("1-4 list overflow");
            }
        }
    }
}

// MoleculeBase::rotateBond (system/molecule.cpp:1099-1180): every atom
// reached from bj without crossing bi turns about the bi -> bj axis
bool Molecule::rotate_bond(int bi, int bj, double angle) {
    if (bi < 0 || bj < 0) return false;
    bool found = false;
    for (const CovAtom &c : atom[bi].cov12) found |= c.i == bj;
    if (!found) return false;
    const Mat R = axis_rot(sub(atom[bj].pos, atom[bi].pos), angle);
    std::vector<int> status(atom.size(), -1), list;
    status[bi] = 2;
    status[bj] = 1;
    list.push_back(bj);
    for (size_t cur = 0; cur < list.size(); ++cur) {
        const int a = list[cur];
        if (status[a] >= 2) continue;
        const Vec o = atom[bi].pos;
        const Vec v = mulmat(sub(atom[a].pos, o), R);
        atom[a].pos = {v.x + o.x, v.y + o.y, v.z + o.z};
        status[a] = 2;
        for (const CovAtom &c : atom[a].cov12) {
            if (c.i == a || status[c.i] >= 1) continue;
            status[c.i] = 1;
            list.push_back(c.i);
        }
    }
    return true;
}

void Molecule::phipsi(int ir, double &phi, double &psi) const {
    const int nr = (int)res.size();
    phi = 0.0;
    psi = 0.0;
    if (ir < nr && ir > 0 && res[ir - 1].iC >= 0 && res[ir].iN >= 0 && res[ir].iCA >= 0 && res[ir].iC >= 0)
        phi = torsion_angle(atom[res[ir - 1].iC].pos, atom[res[ir].iN].pos, atom[res[ir].iCA].pos, atom[res[ir].iC].pos);
    if (ir < nr - 1 && ir >= 0 && res[ir].iN >= 0 && res[ir].iCA >= 0 && res[ir].iC >= 0 && res[ir + 1].iN >= 0)
        psi = torsion_angle(atom[res[ir].iN].pos, atom[res[ir].iCA].pos, atom[res[ir].iC].pos, atom[res[ir + 1].iN].pos);
}

void Molecule::omega(int ir, double &om) const {
    om = 0.0;
    if (ir < (int)res.size() && ir > 0 && res[ir - 1].iCA >= 0 && res[ir - 1].iC >= 0 && res[ir].iN >= 0 &&
        res[ir].iCA >= 0)
        om = torsion_angle(atom[res[ir - 1].iCA].pos, atom[res[ir - 1].iC].pos, atom[res[ir].iN].pos,
                           atom[res[ir].iCA].pos);
}

// MoleculeBase::setResiduePhiPsi / setResidueOmega (system/molecule.cpp:702-721)
void Molecule::set_phipsi(int ir, double nphi, double npsi) {
    double fphi, fpsi;
    phipsi(ir, fphi, fpsi);
    rotate_bond(res[ir].iN, res[ir].iCA, nphi - fphi);
    rotate_bond(res[ir].iCA, res[ir].iC, npsi - fpsi);
}

void Molecule::set_omega(int ir, double nomega) {
    double f;
    omega(ir, f);
    rotate_bond(res[ir - 1].iC, res[ir].iN, nomega - f);
}

// FF_Bonded::setup: assembleBondList / assembleAngleList /
// assembleTorsionList / assembleImproperList (forcefields/ffbonded.cpp:
// 13-460), with the equilibrium 1-2 and 1-3 distances recorded on the atoms
void Molecule::bonded() {
    const ParamSet &P = *ps;
    auto type_of = [&](int a) { return atom[a].tpl->fftype; };
    auto dist = [&](int a, int b) { return mag(sub(atom[a].pos, atom[b].pos)); };
    bond_ij.clear(); bond_lk.clear();
    std::vector<double> bl;
    if (P.autobonds) {
        for (size_t i = 0; i < atom.size(); ++i)
            for (const CovAtom &c : atom[i].cov12) {
                if ((int)i < c.i) continue;
                const int t[2] = {type_of((int)i), type_of(c.i)};
                const int k = best_match(P.bonds.size(), [&](size_t q) {
                    const int pat[2] = {P.bonds[q].i, P.bonds[q].j};
                    return match_count(pat, t, 2);
                });
                double l, kk;
                if (k >= 0) { kk = P.bonds[k].k; l = P.bonds[k].length; }
                else { kk = 360 * kKcal2J / kNa; l = dist((int)i, c.i); }
                bond_ij.push_back((int)i);
                bond_ij.push_back(c.i);
                bond_lk.push_back(l);
                bond_lk.push_back(kk);
            }
        for (size_t i = 0; i < atom.size(); ++i)
            for (CovAtom &c : atom[i].cov12)
                for (size_t b = 0; b < bond_lk.size() / 2; ++b)
                    if ((bond_ij[2 * b] == (int)i && bond_ij[2 * b + 1] == c.i) ||
                        (bond_ij[2 * b + 1] == (int)i && bond_ij[2 * b] == c.i)) {
                        c.d = bond_lk[2 * b];
                        break;
                    }
    }
    angle_iaj.clear(); angle_tk.clear();
    if (P.autoangles) {
        for (size_t i = 0; i < atom.size(); ++i)
            for (const CovAtom &c : atom[i].cov13) {
                if ((int)i < c.i) continue;
                const int ia = c.i2, ij = c.i;
                const int t[3] = {type_of((int)i), type_of(ia), type_of(ij)};
                const int k = best_match(P.angles.size(), [&](size_t q) {
                    const int pat[3] = {P.angles[q].i, P.angles[q].a, P.angles[q].j};
                    return match_count(pat, t, 3);
                });
                double th, kk;
                if (k >= 0) { kk = P.angles[k].k; th = P.angles[k].theta0; }
                else {
                    kk = 0.03 * kKcal2J / kNa;
                    th = angle_with(sub(atom[i].pos, atom[ia].pos), sub(atom[ij].pos, atom[ia].pos));
                }
                angle_iaj.push_back((int)i);
                angle_iaj.push_back(ia);
                angle_iaj.push_back(ij);
                angle_tk.push_back(th);
                angle_tk.push_back(kk);
            }
        auto d12 = [&](int a, int b) {
            for (const CovAtom &c : atom[a].cov12)
                if (c.i == b) return c.d;
            throw std::runtime_error("angle without its bonds");
        };
        for (size_t i = 0; i < atom.size(); ++i)
            for (CovAtom &c : atom[i].cov13)
                for (size_t q = 0; q < angle_tk.size() / 2; ++q) {
                    const int ai = angle_iaj[3 * q], aa = angle_iaj[3 * q + 1], aj = angle_iaj[3 * q + 2];
                    if (aa != c.i2 || !((ai == (int)i && aj == c.i) || (aj == (int)i && ai == c.i))) continue;
                    const double db1 = d12(aa, ai), db2 = d12(aa, aj);
                    c.d = std::sqrt(db1 * db1 + db2 * db2 - 2 * db1 * db2 * std::cos(angle_tk[2 * q]));
                    break;
                }
    }
    auto push_dihedral = [](const TorsionType &T, std::vector<int> &terms, std::vector<double> &vn,
                            std::vector<double> &nn, std::vector<double> &gm) {
        terms.push_back(T.terms);
        for (int m = 0; m < 4; ++m) {
            const bool used = m < T.terms;
            vn.push_back(used ? T.Vn[m] : 0.0);
            nn.push_back(used ? T.n[m] : 0.0);
            gm.push_back(used ? T.gamma[m] : 0.0);
        }
    };
    tors_iabj.clear(); tors_terms.clear(); tors_vn.clear(); tors_n.clear(); tors_gamma.clear();
    if (P.autotorsions) {
        for (size_t i = 0; i < atom.size(); ++i)
            for (const CovAtom &c : atom[i].cov14) {
                if ((int)i < c.i) continue;
                const int q4[4] = {(int)i, c.i3, c.i2, c.i};
                const int t[4] = {type_of(q4[0]), type_of(q4[1]), type_of(q4[2]), type_of(q4[3])};
                const int k = best_match(P.torsions.size(), [&](size_t q) {
                    if (P.torsions[q].kind != 0) return -1;
                    const int pat[4] = {P.torsions[q].i, P.torsions[q].a, P.torsions[q].b, P.torsions[q].j};
                    return match_count(pat, t, 4);
                });
                if (k < 0) continue; // PD drops torsions without parameters
                tors_iabj.insert(tors_iabj.end(), q4, q4 + 4);
                push_dihedral(P.torsions[k], tors_terms, tors_vn, tors_n, tors_gamma);
            }
    }
    imp_iabj.clear(); imp_terms.clear(); imp_vn.clear(); imp_n.clear(); imp_gamma.clear();
    for (size_t ir = 0; ir < res.size(); ++ir) {
        for (const DihedralDef &d : ps->mols[res[ir].mol].improper) {
            int q4[4];
            bool ok = true;
            for (int k = 0; k < 4; ++k) {
                q4[k] = by_raw((int)ir + d.ro[k], d.an[k]);
                ok &= q4[k] >= 0;
            }
            if (!ok) continue;
            const int t[4] = {type_of(q4[0]), type_of(q4[1]), type_of(q4[2]), type_of(q4[3])};
            const int k = best_match(P.torsions.size(), [&](size_t q) {
                if (P.torsions[q].kind != 1) return -1;
                const int pat[4] = {P.torsions[q].i, P.torsions[q].a, P.torsions[q].b, P.torsions[q].j};
                return match_count(pat, t, 4);
            });
            if (k < 0) continue;
            imp_iabj.insert(imp_iabj.end(), q4, q4 + 4);
            push_dihedral(P.torsions[k], imp_terms, imp_vn, imp_n, imp_gamma);
        }
    }
}

// --- sequence grammar -----------------------------------------------------
struct ResToken {
    std::string name; // long name after alias lookup
    std::string fix;  // terminal code
    bool post = false;
    std::string full() const { return post ? name + fix : fix + name; }
};

struct Grammar {
    const ParamSet &ps;
    std::vector<ResToken> out;

    void add(const std::string &tok) {
        if (tok.empty()) throw std::runtime_error("empty residue token");
        const bool start = tok[0] == '*';
        if (start && tok.size() == 1) throw std::runtime_error("residue name '*' is invalid");
        const bool end = tok.size() > 1 && tok[tok.size() - 1] == '*';
        if (start && end) throw std::runtime_error("a single residue cannot start and end the polymer");
        if (start && !out.empty()) throw std::runtime_error("polymer start '*' on a non-terminal residue");
        std::string name = tok;
        while (!name.empty() && name[0] == '*') name.erase(0, 1);
        while (!name.empty() && name[name.size() - 1] == '*') name.erase(name.size() - 1);
        // alias lookup (FFParamSet::lookupAlias): the molecule's own name
        int k = ps.find_mol(name);
        if (k >= 0) name = ps.mols[k].name;
        else throw std::runtime_error("unknown residue '" + name + "'");
        ResToken r;
        r.name = name;
        out.push_back(r);
        if (start || end) {
            bool found = false;
            for (const ResidueClass &c : ps.classes) {
                for (const std::string &m : c.members)
                    if (m == name) {
                        out.back().fix = start ? c.start : c.end;
                        out.back().post = start ? c.start_post : c.end_post;
                        found = true;
                    }
                if (found) break;
                for (const std::string &m : c.caps)
                    if (m == name) found = true; // caps take no terminal code
                if (found) break;
            }
            if (!found) throw std::runtime_error("no polymer class for terminal residue '" + name + "'");
            if (end)
                for (size_t i = 1; i + 1 < out.size(); ++i)
                    if (!out[i].fix.empty()) throw std::runtime_error("multiple termini definitions");
        }
    }

    static int repeat(const std::string &s, size_t &p) {
        std::string digits;
        while (true) {
            ++p;
            if (p >= s.size()) throw std::runtime_error("unexpected end inside '[]'");
            if (s[p] == ']') {
                if (digits.empty()) throw std::runtime_error("no number within '[]'");
                const int n = std::atoi(digits.c_str());
                if (n <= 0) throw std::runtime_error("non-positive count within '[]'");
                return n;
            }
            if (!std::isdigit((unsigned char)s[p])) throw std::runtime_error("non-numeric character within '[]'");
            digits += s[p];
        }
    }

    // BioSequence::append(const std::string&) (sequence/sequence.cpp:304-452)
    void parse(const std::string &s) {
        std::string tok;
        bool bracket = false;
        for (size_t p = 0;; ++p) {
            if (p >= s.size()) {
                if (bracket) throw std::runtime_error("closing ')' expected");
                if (!tok.empty()) add(tok);
                return;
            }
            const char ch = s[p];
            if (ch == '>') throw std::runtime_error("closing '>' unexpected");
            if (ch == '<') {
                if (!tok.empty()) throw std::runtime_error("missing '-' before '<'");
                // the matching '>' (nesting counted)
                size_t depth = 0, q = p + 1;
                for (; q < s.size(); ++q) {
                    if (s[q] == '<') ++depth;
                    else if (s[q] == '>') {
                        if (depth == 0) break;
                        --depth;
                    }
                }
                if (q >= s.size()) throw std::runtime_error("closing '>' expected");
                std::string inner = s.substr(p + 1, q - p - 1);
                if (bracket) inner = "(" + inner + ")";
                p = q;
                int count = 1;
                if (p + 1 < s.size() && s[p + 1] == '[') { ++p; count = repeat(s, p); }
                for (int k = 0; k < count; ++k) parse(inner);
                continue;
            }
            if (ch == '[') {
                if (tok.empty()) throw std::runtime_error("identifier expected before '[]'");
                const int count = repeat(s, p);
                for (int k = 0; k < count; ++k) add(tok);
                tok.clear();
                continue;
            }
            if (bracket) {
                if (ch == '(') throw std::runtime_error("opening '(' unexpected");
                if (ch == '*') throw std::runtime_error("terminus '*' inside a bracket");
                if (ch == ')') { bracket = false; continue; }
                if (!std::isalnum((unsigned char)ch) && ch != '_') throw std::runtime_error("illegal character");
                const std::string one(1, ch);
                int count = 1;
                if (p + 1 < s.size() && s[p + 1] == '[') { ++p; count = repeat(s, p); }
                for (int k = 0; k < count; ++k) add(one);
                continue;
            }
            if (ch == ')') throw std::runtime_error("closing ')' unexpected");
            if (ch == '-') {
                if (!tok.empty()) { add(tok); tok.clear(); }
                continue;
            }
            if (tok.empty() && ch == '(') { bracket = true; continue; }
            const char u = (char)std::toupper((unsigned char)ch);
            if (!std::isalnum((unsigned char)u) && u != '*' && u != '_') throw std::runtime_error("illegal character");
            tok += u;
        }
    }
};

// GeneratorCore + NewProtein (system/genpolymer.cpp:12-164,256-270)
std::unique_ptr<Molecule> build_protein(const ParamSet &ps, const std::string &seq) {
    Grammar g{ps, {}};
    g.parse(seq);
    if (g.out.empty()) throw std::runtime_error("empty sequence");
    std::unique_ptr<Molecule> M(new Molecule);
    M->ps = &ps;
    std::vector<int> def;
    for (const ResToken &r : g.out) {
        const int k = ps.find_mol(r.full());
        if (k < 0) throw std::runtime_error("residue '" + r.full() + "' unknown");
        def.push_back(k);
    }
    for (size_t ir = 0; ir < def.size(); ++ir)
        for (const TemplateAtom &ta : ps.mols[def[ir]].atom) {
            Particle p;
            p.tpl = &ta;
            p.mol = def[ir];
            p.ir = (int)ir;
            p.pos = ta.pos;
            M->atom.push_back(p);
        }
    M->residues();
    // polymerise from the last residue back: residue ir (and everything
    // after it) is moved so that its '-' link point lands on the previous
    // residue's link atom, turned onto the previous residue's '+' link
    for (int ir = (int)def.size() - 1; ir > 0; --ir) {
        const MolDef &cur = ps.mols[def[ir]], &prev = ps.mols[def[ir - 1]];
        const int back = M->by_pdb(ir - 1, cur.backname), frwd = M->by_pdb(ir, prev.frwdname);
        if (back < 0 || frwd < 0) throw std::runtime_error("polymer link atoms not found");
        // superimpose (two-point form, maths/maths.h:323-339)
        const Vec toa = sub(prev.frwdpos, M->atom[back].pos);
        const Vec oa = sub(M->atom[frwd].pos, cur.backpos);
        const Mat R = axis_rot(cross(toa, oa), -angle_with(toa, oa));
        const Vec subt = cur.backpos, addt = M->atom[back].pos;
        for (size_t i = M->res[ir].ifirst; i < M->atom.size(); ++i) {
            const Vec v = mulmat(sub(M->atom[i].pos, subt), R);
            M->atom[i].pos = {v.x + addt.x, v.y + addt.y, v.z + addt.z};
        }
    }
    M->covalent();
    const int nr = (int)M->res.size();
    // setProteinAlphaHelix, then NewProtein's extended chain
    for (int ir = 0; ir < nr; ++ir) M->set_phipsi(ir, -57.0 * kPIOver180, -47.0 * kPIOver180);
    for (int ir = 1; ir < nr; ++ir) M->set_omega(ir, kPI);
    for (int ir = 0; ir < nr; ++ir) M->set_phipsi(ir, kPI, kPI);
    for (int ir = 1; ir < nr; ++ir) M->set_omega(ir, kPI);
    for (Particle &p : M->atom) p.ref = p.pos;
    M->bonded();
    return M;
}

thread_local std::string g_err;

} // namespace

struct pdgpu_ff {
    ParamSet ps;
};
struct pdgpu_molecule {
    std::unique_ptr<Molecule> m;
};

extern "C" const char *pdgpu_prep_error(void) { return g_err.c_str(); }

extern "C" int pdgpu_ff_read(const char *path, pdgpu_ff **out) {
    if (!path || !out) return PDGPU_ERR_INVALID;
    *out = nullptr;
    try {
        std::unique_ptr<pdgpu_ff> f(new pdgpu_ff);
        f->ps.read(path, 0);
        if (!f->ps.error.empty()) { g_err = f->ps.error; return PDGPU_ERR_INVALID; }
        *out = f.release();
        return PDGPU_OK;
    } catch (const std::exception &e) {
        g_err = e.what();
        return PDGPU_ERR_INVALID;
    }
}

extern "C" void pdgpu_ff_free(pdgpu_ff *ff) { delete ff; }

extern "C" int pdgpu_ff_info(const pdgpu_ff *ff, double *vdw14, double *elec14, double *gb_p) {
    if (!ff) return PDGPU_ERR_INVALID;
    if (vdw14) *vdw14 = ff->ps.vdw14;
    if (elec14) *elec14 = ff->ps.elec14;
    if (gb_p) std::memcpy(gb_p, ff->ps.gb_p, sizeof(ff->ps.gb_p));
    return PDGPU_OK;
}

extern "C" int pdgpu_peptide_build(const pdgpu_ff *ff, const char *sequence, pdgpu_molecule **out) {
    if (!ff || !sequence || !out) return PDGPU_ERR_INVALID;
    *out = nullptr;
    try {
        std::unique_ptr<pdgpu_molecule> m(new pdgpu_molecule);
        m->m = build_protein(ff->ps, sequence);
        for (const Particle &p : m->m->atom)
            if (ff->ps.gb_radius.empty() || ff->ps.gb_radius[p.tpl->fftype] < 0) {
                g_err = "no Still radius for atom type " + p.tpl->type_name;
                return PDGPU_ERR_INVALID;
            }
        *out = m.release();
        return PDGPU_OK;
    } catch (const std::exception &e) {
        g_err = e.what();
        return PDGPU_ERR_INVALID;
    }
}

extern "C" void pdgpu_molecule_free(pdgpu_molecule *m) { delete m; }

extern "C" int pdgpu_molecule_counts(const pdgpu_molecule *m, int32_t *counts) {
    if (!m || !counts) return PDGPU_ERR_INVALID;
    const Molecule &M = *m->m;
    int64_t n12 = 0, n13 = 0, n14 = 0;
    for (size_t i = 0; i < M.atom.size(); ++i) {
        for (const CovAtom &c : M.atom[i].cov12) n12 += c.i > (int)i;
        for (const CovAtom &c : M.atom[i].cov13) n13 += c.i > (int)i;
        for (const CovAtom &c : M.atom[i].cov14) n14 += c.i > (int)i;
    }
    counts[0] = (int32_t)M.atom.size();
    counts[1] = (int32_t)M.res.size();
    counts[2] = (int32_t)n12;
    counts[3] = (int32_t)n13;
    counts[4] = (int32_t)n14;
    counts[5] = (int32_t)(M.bond_lk.size() / 2);
    counts[6] = (int32_t)(M.angle_tk.size() / 2);
    counts[7] = (int32_t)M.tors_terms.size();
    counts[8] = (int32_t)M.imp_terms.size();
    return PDGPU_OK;
}

extern "C" int pdgpu_molecule_atoms(const pdgpu_molecule *m, double *xyz, double *charge, double *radius,
                                    double *epsilon, double *mass, double *gb_radius, int32_t *resid,
                                    uint8_t *is_hydrogen) {
    if (!m) return PDGPU_ERR_INVALID;
    const Molecule &M = *m->m;
    for (size_t i = 0; i < M.atom.size(); ++i) {
        const Particle &p = M.atom[i];
        const TemplateAtom &t = *p.tpl;
        if (xyz) { xyz[3 * i] = p.pos.x; xyz[3 * i + 1] = p.pos.y; xyz[3 * i + 2] = p.pos.z; }
        if (charge) charge[i] = t.charge;
        if (radius) radius[i] = t.radius;
        if (epsilon) epsilon[i] = t.epsilon;
        if (mass) mass[i] = t.mass;
        if (gb_radius) gb_radius[i] = M.ps->gb_radius[t.fftype];
        if (resid) resid[i] = p.ir;
        // hackAtomFlags (forcefields/ffparam.cpp:739-797): hydrogen by the
        // first letter of the atom type's name
        if (is_hydrogen) is_hydrogen[i] = M.ps->types[t.fftype].name[0] == 'H';
    }
    return PDGPU_OK;
}

extern "C" int pdgpu_molecule_atom_names(const pdgpu_molecule *m, int32_t i, const char **pdbname,
                                         const char **type_name, const char **rawname, const char **parentname) {
    if (!m || i < 0 || i >= (int32_t)m->m->atom.size()) return PDGPU_ERR_INVALID;
    const Particle &p = m->m->atom[i];
    if (pdbname) *pdbname = p.tpl->pdbname.c_str();
    if (type_name) *type_name = p.tpl->type_name.c_str();
    if (rawname) *rawname = p.tpl->rawname.c_str();
    if (parentname) *parentname = m->m->ps->mols[p.mol].name.c_str();
    return PDGPU_OK;
}

extern "C" int pdgpu_molecule_pairs(const pdgpu_molecule *m, int32_t order, int32_t *a, int32_t *b, double *d) {
    if (!m || order < 2 || order > 4) return PDGPU_ERR_INVALID;
    const Molecule &M = *m->m;
    int64_t k = 0;
    for (size_t i = 0; i < M.atom.size(); ++i) {
        const std::vector<CovAtom> &L = order == 2 ? M.atom[i].cov12 : order == 3 ? M.atom[i].cov13 : M.atom[i].cov14;
        for (const CovAtom &c : L) {
            if (c.i <= (int)i) continue;
            if (a) a[k] = (int32_t)i;
            if (b) b[k] = c.i;
            if (d) d[k] = c.d;
            ++k;
        }
    }
    return PDGPU_OK;
}

extern "C" int pdgpu_molecule_cov12(const pdgpu_molecule *m, int32_t *cov12, int32_t *n_cov12) {
    if (!m) return PDGPU_ERR_INVALID;
    const Molecule &M = *m->m;
    for (size_t i = 0; i < M.atom.size(); ++i) {
        const std::vector<CovAtom> &L = M.atom[i].cov12;
        if (n_cov12) n_cov12[i] = (int32_t)L.size();
        for (int k = 0; k < 6 && cov12; ++k) cov12[6 * i + k] = k < (int)L.size() ? L[k].i : -1;
    }
    return PDGPU_OK;
}

extern "C" int pdgpu_molecule_backbone(const pdgpu_molecule *m, int32_t *n_ca_c) {
    if (!m || !n_ca_c) return PDGPU_ERR_INVALID;
    const Molecule &M = *m->m;
    for (size_t r = 0; r < M.res.size(); ++r) {
        n_ca_c[3 * r] = M.res[r].iN;
        n_ca_c[3 * r + 1] = M.res[r].iCA;
        n_ca_c[3 * r + 2] = M.res[r].iC;
    }
    return PDGPU_OK;
}

extern "C" int pdgpu_molecule_set_phipsi(pdgpu_molecule *m, int32_t ir, double phi, double psi, double omega) {
    if (!m || ir < 0 || ir >= (int32_t)m->m->res.size()) return PDGPU_ERR_INVALID;
    m->m->set_phipsi(ir, phi, psi);
    if (ir > 0 && std::isfinite(omega)) m->m->set_omega(ir, omega);
    return PDGPU_OK;
}

extern "C" int pdgpu_molecule_get_phipsi(const pdgpu_molecule *m, int32_t ir, double *phi, double *psi,
                                         double *omega) {
    if (!m || ir < 0 || ir >= (int32_t)m->m->res.size()) return PDGPU_ERR_INVALID;
    double a, b, c;
    m->m->phipsi(ir, a, b);
    m->m->omega(ir, c);
    if (phi) *phi = a;
    if (psi) *psi = b;
    if (omega) *omega = c;
    return PDGPU_OK;
}

extern "C" int pdgpu_molecule_bonded(const pdgpu_molecule *m, pdgpu_bonded_out *out) {
    if (!m || !out) return PDGPU_ERR_INVALID;
    const Molecule &M = *m->m;
    auto put_i = [](int32_t *dst, const std::vector<int> &src) {
        if (dst) for (size_t k = 0; k < src.size(); ++k) dst[k] = src[k];
    };
    auto put_d = [](double *dst, const std::vector<double> &src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(double));
    };
    out->nbonds = (int32_t)(M.bond_lk.size() / 2);
    out->nangles = (int32_t)(M.angle_tk.size() / 2);
    out->ntorsions = (int32_t)M.tors_terms.size();
    out->nimpropers = (int32_t)M.imp_terms.size();
    put_i(out->bond_ij, M.bond_ij);
    put_d(out->bond_lk, M.bond_lk);
    put_i(out->angle_iaj, M.angle_iaj);
    put_d(out->angle_tk, M.angle_tk);
    put_i(out->torsion_iabj, M.tors_iabj);
    put_i(out->torsion_terms, M.tors_terms);
    put_d(out->torsion_vn, M.tors_vn);
    put_d(out->torsion_n, M.tors_n);
    put_d(out->torsion_gamma, M.tors_gamma);
    put_i(out->improper_iabj, M.imp_iabj);
    put_i(out->improper_terms, M.imp_terms);
    put_d(out->improper_vn, M.imp_vn);
    put_d(out->improper_n, M.imp_n);
    put_d(out->improper_gamma, M.imp_gamma);
    return PDGPU_OK;
}
