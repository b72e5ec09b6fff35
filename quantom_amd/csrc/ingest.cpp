// ingest.cpp — system ingest behind include/lima_ingest.h (host code only).
//
// Three stages with the reference's text files between them:
//   stage 1  CHARMM .prm merge           -> LIMA_ffnonbonded.txt, LIMA_ffbonded.txt
//   stage 2  per-system type filter      -> LIMA_ffnonbonded_filtered.txt, LIMA_ffbonded_filtered.txt
//   stage 3  molecule + box build        -> lima_system arrays owned by a lima_ingest handle
//
// The arithmetic is FP32 in the reference's operation order (built without FMA
// contraction), so that every parameter and coordinate equals the reference's bit for
// bit; the text formats are the reference's so that files pass between the two.
// Where the reference reads past its own rules (undefined behaviour) this file takes
// one defined reading, named where it happens.
#include "../../include/lima_ingest.h"

#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

struct IngestError : std::runtime_error {
    int code;
    IngestError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

thread_local std::string g_last_error;

int fail(const std::exception& e) {
    g_last_error = e.what();
    if (auto* ie = dynamic_cast<const IngestError*>(&e)) return ie->code;
    return LIMA_E_ARG;
}

[[noreturn]] void bad(const std::string& m) { throw IngestError(LIMA_E_ARG, m); }

using Row = std::vector<std::string>;

// ------------------------------------------------------------------ text and numbers

// Lines as std::getline yields them: split at '\n', no empty line after a final '\n'.
std::vector<std::string> file_lines(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) bad("cannot open " + path);
    std::string text;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    std::fclose(f);
    std::vector<std::string> lines;
    size_t at = 0;
    while (at < text.size()) {
        size_t e = text.find('\n', at);
        if (e == std::string::npos) e = text.size();
        lines.emplace_back(text, at, e - at);
        at = e + 1;
    }
    return lines;
}

void write_text(const std::string& path, const std::string& text) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) bad("cannot write " + path);
    size_t w = std::fwrite(text.data(), 1, text.size(), f);
    if (std::fclose(f) != 0 || w != text.size()) bad("short write to " + path);
}

// Non-empty pieces of `s` between any of the separator characters.
Row tokens(const std::string& s, const char* seps) {
    Row out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && std::strchr(seps, s[i])) i++;
        size_t j = i;
        while (j < s.size() && !std::strchr(seps, s[j])) j++;
        if (j > i) out.emplace_back(s, i, j - i);
        i = j;
    }
    return out;
}

// The parameter programs' reader (Filehandling.h:29-110): a line is skipped when empty
// or when its first non-blank character is one of `skip`; CHARMM files split at blanks
// and tabs, the programs' own files at blanks and ';'.
std::vector<Row> param_rows(const std::string& path, const char* skip, bool own_format) {
    std::vector<Row> rows;
    for (const std::string& line : file_lines(path)) {
        if (line.empty()) continue;
        char first = ' ';
        for (char c : line)
            if (c != ' ' && c != '\t') { first = c; break; }
        if (std::strchr(skip, first) && first != ' ') continue;
        rows.push_back(tokens(line, own_format ? " ;" : " \t"));
    }
    return rows;
}

// The engine's reader (Filehandling.cpp:7-66): skips empty lines and lines starting with
// ';' or '/', splits at blanks and ';'.
std::vector<Row> engine_rows(const std::string& path) {
    std::vector<Row> rows;
    for (const std::string& line : file_lines(path)) {
        if (line.empty() || line[0] == ';' || line[0] == '/') continue;
        rows.push_back(tokens(line, " ;"));
    }
    return rows;
}

// The molecule builder's split of .gro/.top lines (CompoundBuilder.cpp:324-360, 415-436):
// single blanks only; empty words and lone brackets dropped.
Row gro_words(const std::string& line) {
    Row out;
    for (std::string& w : tokens(line, " "))
        if (w != "[" && w != "]") out.push_back(std::move(w));
    return out;
}

float parse_f(const std::string& s) {                  // std::stof
    const char* b = s.c_str();
    char* e = nullptr;
    float v = std::strtof(b, &e);
    if (e == b) bad("not a number: '" + s + "'");
    return v;
}

float parse_f_via_double(const std::string& s) {       // Float3(stod(s), ...)
    const char* b = s.c_str();
    char* e = nullptr;
    double v = std::strtod(b, &e);
    if (e == b) bad("not a number: '" + s + "'");
    return (float)v;
}

int parse_i(const std::string& s) {                     // std::stoi
    const char* b = s.c_str();
    char* e = nullptr;
    errno = 0;
    long v = std::strtol(b, &e, 10);
    if (e == b || errno == ERANGE || v > 2147483647L || v < -2147483647L - 1) bad("not an integer: '" + s + "'");
    return (int)v;
}

std::string fmt(float x) {                              // std::to_string(float)
    char b[512];
    std::snprintf(b, sizeof b, "%f", (double)x);
    return b;
}

// ------------------------------------------------------------------ type-name order

// The parameter tables key a term by its type names in one canonical orientation.
// Pair order (ForcefieldTypes.h:399-423, 547-562): at the first differing character the
// smaller one goes first; when one name is a prefix of the other, the shorter goes first.
void order_outer(std::string& a, std::string& b) {
    for (size_t p = 0; p < a.size() && p < b.size(); p++) {
        if (a[p] < b[p]) return;
        if (a[p] > b[p]) { std::swap(a, b); return; }
    }
    if (a.size() > b.size()) std::swap(a, b);
}

// FTHelpers::isSorted (ForcefieldTypes.h:42-56).  The reference falls off the end without
// a return value when the left name is a prefix of the right one; that case reads "sorted".
bool in_order(const std::string& a, const std::string& b) {
    for (size_t p = 0; p < a.size() && p < b.size(); p++) {
        if (a[p] < b[p]) return true;
        if (a[p] > b[p]) return false;
    }
    return !(a.size() > b.size());
}

template <size_t N> using Names = std::array<std::string, N>;

void canonical(Names<2>& t) { order_outer(t[0], t[1]); }
void canonical(Names<3>& t) { order_outer(t[0], t[2]); }   // the centre stays in the centre
void canonical(Names<4>& t) {                              // Dihedraltype::sort (ForcefieldTypes.h:706-721)
    bool flip = t[0] != t[3] ? !in_order(t[0], t[3]) : !in_order(t[1], t[2]);
    if (flip) { std::swap(t[0], t[3]); std::swap(t[1], t[2]); }
}

// ------------------------------------------------------------------ stage 1: CHARMM merge

struct AtomType {
    std::string name;
    float mass = -1.f, sigma = -1.f, epsilon = -1.f;   // [g/mol], [nm], [J/mol]
    int local_id = 0;
};

template <size_t N> struct TermType {
    Names<N> names;
    float a = 0.f, b = 0.f;   // bonds: b0, kb; angles: theta0, ktheta; dihedrals: phi0, kphi
    int n = 0;                // dihedral multiplicity
};

struct ParamTables {
    std::vector<AtomType> atoms;
    std::vector<TermType<2>> bonds;
    std::vector<TermType<3>> angles;
    std::vector<TermType<4>> dihedrals;
};

float deg_to_rad(float deg) { return deg * 2 * 3.1415f / 360.f; }   // ForcefieldMerger.h:127, 162

enum class PrmSection { none, atoms, bonds, angles, dihedrals, nonbonded, improper, cmap };

PrmSection prm_section(const std::string& w, PrmSection cur) {        // ForcefieldMerger.h:184-201
    static const std::map<std::string, PrmSection> keys = {
        {"ATOMS", PrmSection::atoms}, {"BONDS", PrmSection::bonds}, {"ANGLES", PrmSection::angles},
        {"DIHEDRALS", PrmSection::dihedrals}, {"NONBONDED", PrmSection::nonbonded},
        {"IMPROPER", PrmSection::improper}, {"CMAP", PrmSection::cmap}};
    auto it = keys.find(w);
    return it == keys.end() ? cur : it->second;
}

// ForcefieldMerger::mergeForcefields (ForcefieldMerger.h:10-176).  Sections are known only
// by their keyword rows, so NBFIX rows (which start no section) are read as NONBONDED rows.
ParamTables merge_prm(const std::vector<std::string>& paths) {
    ParamTables t;
    std::unordered_map<std::string, size_t> atom_at;
    std::set<Names<2>> seen2;
    std::set<Names<3>> seen3;
    std::set<Names<4>> seen4;
    for (const std::string& path : paths) {
        PrmSection sec = PrmSection::none;
        for (const Row& r : param_rows(path, ";#!", false)) {
            if (r.empty()) continue;
            sec = prm_section(r[0], sec);
            switch (sec) {
            case PrmSection::atoms:                          // MASS <id> <type> <mass>
                if (r.size() < 4 || atom_at.count(r[2])) break;
                atom_at[r[2]] = t.atoms.size();
                t.atoms.push_back({r[2], parse_f(r[3])});
                break;
            case PrmSection::nonbonded: {                    // <type> <ignored> <epsilon> <Rmin/2>
                if (r.size() < 4) break;
                auto it = atom_at.find(r[0]);
                if (it == atom_at.end()) break;
                AtomType& a = t.atoms[it->second];
                a.epsilon = std::fabs(parse_f(r[2]) * 4184.f);
                a.sigma = parse_f(r[3]) * 2.f / 10.f;
                break;
            }
            case PrmSection::bonds: {                        // <t1> <t2> <Kb> <b0>
                if (r.size() < 4) break;
                TermType<2> b{{r[0], r[1]}, parse_f(r[3]) * 0.1f, parse_f(r[2]) * 4183.f * 100.f};
                canonical(b.names);
                if (seen2.insert(b.names).second) t.bonds.push_back(b);
                break;
            }
            case PrmSection::angles: {                       // <t1> <t2> <t3> <Ktheta> <theta0>
                if (r.size() < 5) break;
                TermType<3> a{{r[0], r[1], r[2]}, deg_to_rad(parse_f(r[4])), parse_f(r[3]) * 4183.f};
                canonical(a.names);
                if (seen3.insert(a.names).second) t.angles.push_back(a);
                break;
            }
            case PrmSection::dihedrals: {                    // <t1..t4> <Kchi> <n> <delta>
                if (r.size() < 7) break;
                TermType<4> d{{r[0], r[1], r[2], r[3]}, deg_to_rad(parse_f(r[6])), parse_f(r[4]) * 4183.f,
                              parse_i(r[5])};
                canonical(d.names);
                if (seen4.insert(d.names).second) t.dihedrals.push_back(d);
                break;
            }
            default:
                break;
            }
        }
    }
    return t;
}

// Header lines of the reference's text files (Filehandling.h:117-135).  Readers skip
// them; they are kept so that the files match the reference's byte for byte.
std::string h1(const std::string& s) { return "// ----------------################ " + s + " ################---------------- //\n\n"; }
std::string h2(const std::string& s) { return "// ---------------- " + s + " ---------------- //\n"; }
std::string h3(const std::string& s) { return "// " + s + " //\n"; }
std::string section(const std::string& s) { return "# " + s + "\n"; }
const char* kEndBlock = "\n\n\n\n";

std::string join(std::initializer_list<std::string> parts) {
    std::string out;
    bool first = true;
    for (const std::string& p : parts) {
        if (!first) out += ';';
        out += p;
        first = false;
    }
    return out + "\n";
}

std::string merged_nonbonded_text(const ParamTables& t) {        // Filehandling.h:225-242
    std::string o = h1("Forcefield Non-bonded") + h2("Non-bonded parameters") +
                    h3("{atom_type \t mass [g/mol] \t sigma [nm] \t epsilon [J/mol]}") + section("ff_nonbonded");
    for (const AtomType& a : t.atoms) o += join({a.name, fmt(a.mass), fmt(a.sigma), fmt(a.epsilon)});
    return o + kEndBlock;
}

std::string merged_bonded_text(const ParamTables& t) {           // Filehandling.h:244-286
    std::string o = h1("Forcefield Bonded") + h2("Bondtype parameters") +
                    h3("{atom_types \t b_0 [nm] \t k_b [J/(mol*nm^2)] \t }") + section("ff_bondtypes");
    for (const auto& b : t.bonds) o += join({b.names[0], b.names[1], fmt(b.a), fmt(b.b)});
    o += std::string(kEndBlock) + h2("Angletype parameters") +
         h3("{atom_types \t theta_0 [rad] \t k_theta [J/(mol*rad^2)] \t }") + section("ff_angletypes");
    for (const auto& a : t.angles) o += join({a.names[0], a.names[1], a.names[2], fmt(a.a), fmt(a.b)});
    o += std::string(kEndBlock) + h2("Dihedraltype parameters") +
         h3("{atom_types \t phi_0 [rad] \t k_phi [J/(mol)] \t n}") + section("ff_dihedraltypes");
    for (const auto& d : t.dihedrals)
        o += join({d.names[0], d.names[1], d.names[2], d.names[3], fmt(d.a), fmt(d.b), std::to_string(d.n)});
    return o + kEndBlock;
}

// ------------------------------------------------------------------ stage 2: per-system filter

// The solvent type, always slot 0 (ForcefieldTypes.h:17-19, 177).
const float kWaterMass = (float)(15.999000 + 2 * 1.008000);
const float kWaterSigma = (float)(0.302905564168 + 2 * 0.040001352445);
const float kWaterEpsilon = (0.50208f + 2.f * 0.19246f) * 1000.f;

enum class FtSection { none, nonbonded, bonds, angles, dihedrals, pairs };

FtSection ft_section(const std::string& w, FtSection cur) {      // FTHelpers::setState (ForcefieldTypes.h:84-97)
    if (w == "ff_nonbonded") return FtSection::nonbonded;
    if (w == "ff_bondtypes" || w == "bonds") return FtSection::bonds;
    if (w == "ff_angletypes" || w == "angles") return FtSection::angles;
    if (w == "ff_dihedraltypes" || w == "dihedrals") return FtSection::dihedrals;
    if (w == "pairs") return FtSection::pairs;
    return cur;
}

// The merged files read back (ForcefieldTypes.h:172-204, 426-449, 572-599, 725-750).  The
// angle reader stops at the dihedral section; the others read to the end.
ParamTables read_merged(const std::string& nb_path, const std::string& b_path) {
    ParamTables t;
    t.atoms.push_back({"WATER", kWaterMass, kWaterSigma, kWaterEpsilon});
    FtSection sec = FtSection::none;
    for (const Row& r : param_rows(nb_path, "/", true)) {
        if (r.size() == 2) { sec = ft_section(r[1], sec); continue; }
        if (sec == FtSection::nonbonded) {
            if (r.size() < 4) bad("short non-bonded row in " + nb_path);
            t.atoms.push_back({r[0], parse_f(r[1]), parse_f(r[2]), parse_f(r[3])});
        }
    }
    sec = FtSection::none;
    bool angles_closed = false;
    for (const Row& r : param_rows(b_path, "/", true)) {
        if (r.size() == 2) {
            sec = ft_section(r[1], sec);
            if (sec == FtSection::dihedrals) angles_closed = true;
            continue;
        }
        if (sec == FtSection::bonds) {
            if (r.size() < 4) bad("short bond row in " + b_path);
            TermType<2> b{{r[0], r[1]}, parse_f(r[2]), parse_f(r[3])};
            canonical(b.names);
            t.bonds.push_back(b);
        } else if (sec == FtSection::angles && !angles_closed) {
            if (r.size() < 5) bad("short angle row in " + b_path);
            TermType<3> a{{r[0], r[1], r[2]}, parse_f(r[3]), parse_f(r[4])};
            canonical(a.names);
            t.angles.push_back(a);
        } else if (sec == FtSection::dihedrals) {
            if (r.size() < 7) bad("short dihedral row in " + b_path);
            TermType<4> d{{r[0], r[1], r[2], r[3]}, parse_f(r[4]), parse_f(r[5]), parse_i(r[6])};
            canonical(d.names);
            t.dihedrals.push_back(d);
        }
    }
    return t;
}

// FTHelpers::calcLikeness (ForcefieldTypes.h:29-40): the share of the longer name that
// is a common prefix, summed in FP32 one character at a time.
float likeness(const std::string& a, const std::string& b) {
    float step = 1.f / (float)std::max(a.size(), b.size());
    float l = 0.f;
    for (size_t i = 0; i < std::min(a.size(), b.size()); i++) {
        if (a[i] != b[i]) break;
        l += step;
    }
    return l;
}

// The parameter search (ForcefieldTypes.h:492-510, 646-666, 802-822): the first entry
// with the largest product of per-name likeness; angle and dihedral searches stop at an
// exact product of 1.  The result depends on the type names only, so each distinct
// query is searched once.
template <size_t N> class TermMatcher {
public:
    TermMatcher(const std::vector<TermType<N>>& table, float threshold, bool stop_at_one)
        : table_(table), threshold_(threshold), stop_(stop_at_one) {}
    const TermType<N>& operator()(const Names<N>& q) {
        auto it = memo_.find(q);
        if (it != memo_.end()) return table_[it->second];
        float best = 0.f;
        size_t at = table_.size();
        for (size_t i = 0; i < table_.size(); i++) {
            float l = likeness(q[0], table_[i].names[0]);
            for (size_t k = 1; k < N; k++) l = l * likeness(q[k], table_[i].names[k]);
            if (l > best) { best = l; at = i; }
            if (stop_ && l == 1.f) break;
        }
        if (at == table_.size() || !(best > threshold_)) {
            std::string m = "no parameters for types";
            for (const std::string& s : q) m += " " + s;
            bad(m);
        }
        memo_.emplace(q, at);
        return table_[at];
    }

private:
    const std::vector<TermType<N>>& table_;
    float threshold_;
    bool stop_;
    std::map<Names<N>, size_t> memo_;
};

template <size_t N> struct TopTerm {
    std::array<int, N> ids;
    Names<N> names;
    float a = 0.f, b = 0.f;
    int n = 0;
};

struct Filtered {
    std::vector<std::pair<std::string, std::string>> aliases;   // conf atom name -> type name
    std::vector<AtomType> types;                                 // slot order, water first
    std::vector<std::pair<int, int>> atoms;                      // topology id, type slot
    std::vector<TopTerm<2>> bonds;
    std::vector<TopTerm<3>> angles;
    std::vector<TopTerm<4>> dihedrals;
};

Filtered filter_system(const std::string& conf, const std::string& top, const ParamTables& ff) {
    Filtered f;
    // Names of .gro atom lines as the parameter reader splits them (ForcefieldTypes.h:58-80).
    std::vector<std::string> names;
    std::set<std::string> name_seen;
    for (const Row& r : param_rows(conf, ";#!", false))
        if (r.size() == 6 && name_seen.insert(r[1]).second) names.push_back(r[1]);

    // Each name resolves to the first table type equal to its longest prefix that is a
    // type name; types enter the system's table in order of first use
    // (NB_Atomtype::filterUnusedTypes, ForcefieldTypes.h:216-283).
    std::unordered_map<std::string, size_t> first_of;
    for (size_t i = 0; i < ff.atoms.size(); i++) first_of.emplace(ff.atoms[i].name, i);
    f.types.push_back(ff.atoms[0]);
    f.types[0].local_id = 0;
    std::unordered_map<std::string, int> slot_of{{ff.atoms[0].name, 0}};
    std::unordered_map<std::string, std::string> alias_of;
    for (const std::string& name : names) {
        std::string alias = name;
        while (!first_of.count(alias)) {
            if (alias.size() <= 1) bad("no non-bonded type for atom name '" + name + "'");
            alias.pop_back();
        }
        if (!slot_of.count(alias)) {
            AtomType a = ff.atoms[first_of[alias]];
            a.local_id = (int)f.types.size();
            slot_of[alias] = a.local_id;
            f.types.push_back(a);
        }
        if (alias != name && alias_of.emplace(name, alias).second) f.aliases.emplace_back(name, alias);
    }

    // Topology atoms: [ atoms ] rows up to [ bonds ] (Atom::parseTopolAtoms, ForcefieldTypes.h:310-340).
    const std::vector<Row> top_rows = param_rows(top, ";#!", false);
    std::vector<std::string> bonded_type;        // by topology id - 1
    {
        int state = 0;                           // 0 before, 1 in [ atoms ], 2 done
        for (const Row& r : top_rows) {
            if (r.size() == 3) {
                if (r[1] == "atoms") state = 1;
                else if (r[1] == "bonds") state = 2;
                continue;
            }
            if (state == 2) break;
            if (state == 1) {
                if (r.size() < 5) bad("short [ atoms ] row in " + top);
                int id = parse_i(r[0]);
                auto al = alias_of.find(r[4]);
                const std::string& tname = al == alias_of.end() ? r[4] : al->second;
                auto s = slot_of.find(tname);
                if (s == slot_of.end()) bad("topology atom " + r[0] + " (" + r[4] + ") has no type");
                f.atoms.emplace_back(id, s->second);
                bonded_type.push_back(r[1]);
            }
        }
    }
    auto type_of = [&](int id) -> const std::string& {
        if (id < 1 || id > (int)bonded_type.size()) bad("topology term names atom " + std::to_string(id));
        return bonded_type[id - 1];
    };

    // Bonded terms: every [ bonds ] row, [ angles ] rows, and the first [ dihedrals ]
    // section's five-column rows (ForcefieldTypes.h:451-475, 601-630, 752-783).
    TermMatcher<2> match_bond(ff.bonds, 0.01f, false);
    TermMatcher<3> match_angle(ff.angles, 0.01f, true);
    TermMatcher<4> match_dihedral(ff.dihedrals, 0.001f, true);
    FtSection sec = FtSection::none;
    bool angles_open = true, dih_entered = false, dih_closed = false;
    for (const Row& r : top_rows) {
        if (r.size() == 3 && r[0][0] == '[') {
            sec = ft_section(r[1], sec);
            if (sec == FtSection::dihedrals) angles_open = false;
            if (dih_entered) dih_closed = true;
            continue;
        }
        if (sec == FtSection::bonds) {
            if (r.size() < 2) bad("short [ bonds ] row in " + top);
            TopTerm<2> t{{parse_i(r[0]), parse_i(r[1])}, {}};
            for (size_t k = 0; k < 2; k++) t.names[k] = type_of(t.ids[k]);
            canonical(t.names);
            const TermType<2>& m = match_bond(t.names);
            t.a = m.a; t.b = m.b;
            f.bonds.push_back(t);
        } else if (sec == FtSection::angles && angles_open) {
            if (r.size() < 3) bad("short [ angles ] row in " + top);
            TopTerm<3> t{{parse_i(r[0]), parse_i(r[1]), parse_i(r[2])}, {}};
            for (size_t k = 0; k < 3; k++) t.names[k] = type_of(t.ids[k]);
            canonical(t.names);
            const TermType<3>& m = match_angle(t.names);
            t.a = m.a; t.b = m.b;
            f.angles.push_back(t);
        } else if (sec == FtSection::dihedrals && !dih_closed) {
            if (r.size() != 5) continue;
            TopTerm<4> t{{parse_i(r[0]), parse_i(r[1]), parse_i(r[2]), parse_i(r[3])}, {}};
            for (size_t k = 0; k < 4; k++) t.names[k] = type_of(t.ids[k]);
            canonical(t.names);
            const TermType<4>& m = match_dihedral(t.names);
            t.a = m.a; t.b = m.b; t.n = m.n;
            f.dihedrals.push_back(t);
            dih_entered = true;
        }
    }
    return f;
}

std::string filtered_nonbonded_text(const Filtered& f) {        // Filehandling.h:139-165
    std::string o = "#pragma once\n" + h1("Forcefield Non-bonded") + h2("Mappings") + section("mappings");
    for (const auto& m : f.aliases) o += join({m.first, m.second});
    o += std::string(kEndBlock) + h2("Non-bonded parameters") +
         h3("{atom_type \t type_id \t mass [g/mol] \t sigma [nm] \t epsilon [J/mol]}") + section("ff_nonbonded");
    for (const AtomType& a : f.types)
        o += join({a.name, std::to_string(a.local_id), fmt(a.mass), fmt(a.sigma), fmt(a.epsilon)});
    return o + kEndBlock;
}

std::string filtered_bonded_text(const Filtered& f) {           // Filehandling.h:167-223
    auto ids = [](const auto& t) {
        std::string s;
        for (int i : t.ids) s += std::to_string(i) + ";";
        for (const std::string& n : t.names) s += n + ";";
        return s;
    };
    std::string o = h1("Forcefield Non-bonded") +
                    h3("Atoms {particle id [simulation specific]\tatomtype_id [simulation specific]}") + section("atoms");
    for (const auto& a : f.atoms) o += join({std::to_string(a.first), std::to_string(a.second)});
    o += std::string(kEndBlock) + h2("Bonds") +
         h3("{ID_p1 \t ID_p2 \t Atomtype \t Atomtype \t b_0 [nm] \t k_b [J/(mol * nm^2)]}") + section("bonds");
    for (const auto& t : f.bonds) o += ids(t) + fmt(t.a) + ";" + fmt(t.b) + "\n";
    o += std::string(kEndBlock) + h2("Angles") +
         h3("{Atom-IDs \t Atomtypes \t theta_0 [rad] \t k_theta [J/(mol * rad^2)}") + section("angles");
    for (const auto& t : f.angles) o += ids(t) + fmt(t.a) + ";" + fmt(t.b) + "\n";
    o += std::string(kEndBlock) + h2("Dihedrals") +
         h3("{Atom IDs \t Atomtypes \t phi_0 [rad] \t k_phi [J/(mol * rad^2)] \t n}") + section("dihedrals");
    for (const auto& t : f.dihedrals) o += ids(t) + fmt(t.a) + ";" + fmt(t.b) + ";" + std::to_string(t.n) + "\n";
    return o + kEndBlock;
}

// ------------------------------------------------------------------ stage 3: molecule and box

struct F3 {
    float x = 0.f, y = 0.f, z = 0.f;
};
F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// Float3 has only a double scale (QuantomTypes.cuh:39): the product is FP64, narrowed.
F3 scale(F3 a, double s) { return {(float)(a.x * s), (float)(a.y * s), (float)(a.z * s)}; }
float length(F3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }   // QuantomTypes.cuh:60

F3 centre(const std::vector<F3>& p) {                   // Compound::calcCOM (Bodies.cuh:362-369)
    F3 c;
    const float w = 1.f / (float)p.size();
    for (const F3& q : p) c = c + scale(q, w);
    return c;
}

// The engine-side force field (ForceFieldMaker::buildForcefield, Forcefield.cu:10-218).
struct EngineFF {
    lima_forcefield params{};
    int n_types = 0;
    std::vector<int> type_of;                              // by .gro serial; -1 unknown
    std::map<std::array<int, 2>, std::pair<float, float>> bonds;
    std::map<std::array<int, 3>, std::pair<float, float>> angles;
    std::map<std::array<int, 4>, std::tuple<float, float, int>> dihedrals;
};

EngineFF read_filtered(const std::string& nb_path, const std::string& b_path) {
    EngineFF e;
    for (auto& p : e.params.particle_parameters) p = {-1.f, -1.f, -1.f};   // Forcefield.cuh:18-22
    std::string sec;
    auto section_of = [](const std::string& w, const std::string& cur) {
        static const std::set<std::string> keys = {"ff_nonbonded", "atoms", "bonds", "angles", "dihedrals"};
        return keys.count(w) ? w : cur;
    };
    for (const Row& r : engine_rows(nb_path)) {
        if (r.empty()) continue;
        if (r[0][0] == '#') { if (r.size() > 1) sec = section_of(r[1], sec); continue; }
        if (sec != "ff_nonbonded") continue;
        if (r.size() < 5) bad("short row in " + nb_path);
        if (e.n_types == LIMA_MAX_ATOM_TYPES) throw IngestError(LIMA_E_CAPACITY, "more than 32 atom types");
        lima_particle_params& p = e.params.particle_parameters[e.n_types++];
        p.mass = (float)(parse_f(r[2]) * 1e-3);
        p.sigma = parse_f(r[3]);
        p.epsilon = parse_f(r[4]);
    }
    sec.clear();
    for (const Row& r : engine_rows(b_path)) {
        if (r.empty()) continue;
        if (r[0][0] == '#') { if (r.size() > 1) sec = section_of(r[1], sec); continue; }
        if (sec == "atoms") {
            if (r.size() < 2) bad("short row in " + b_path);
            int id = parse_i(r[0]);
            if (id < 0 || id > 100000000) bad("atom id out of range in " + b_path);
            if ((int)e.type_of.size() <= id) e.type_of.resize(id + 1, -1);
            e.type_of[id] = parse_i(r[1]);
        } else if (sec == "bonds") {
            if (r.size() < 6) bad("short row in " + b_path);
            e.bonds.emplace(std::array<int, 2>{parse_i(r[0]), parse_i(r[1])}, std::make_pair(parse_f(r[4]), parse_f(r[5])));
        } else if (sec == "angles") {
            if (r.size() < 8) bad("short row in " + b_path);
            e.angles.emplace(std::array<int, 3>{parse_i(r[0]), parse_i(r[1]), parse_i(r[2])},
                             std::make_pair(parse_f(r[6]), parse_f(r[7])));
        } else if (sec == "dihedrals") {
            if (r.size() < 11) bad("short row in " + b_path);
            e.dihedrals.emplace(std::array<int, 4>{parse_i(r[0]), parse_i(r[1]), parse_i(r[2]), parse_i(r[3])},
                                std::make_tuple(parse_f(r[8]), std::fabs(parse_f(r[9])), parse_i(r[10])));
        }
    }
    return e;
}

struct GroAtom {
    int serial;
    std::string name, residue;
    int residue_id;
    F3 pos;
};

// CompoundBuilder::parseGRO (CompoundBuilder.cpp:415-479, CompoundBuilder.h:90-106).
std::vector<GroAtom> read_gro(const std::string& path) {
    std::vector<GroAtom> out;
    for (const std::string& line : file_lines(path)) {
        Row w = gro_words(line);
        if (w.empty()) bad("blank line in " + path);
        std::string digits, name;
        for (char c : w[0]) (c > 47 && c < 58 ? digits : name) += c;
        if (w.size() == 5) {      // a water atom name fused with a five-digit serial
            Row v{w[0], w[1][0] == 'O' ? "OW" : "HW", w[1].substr(w[1][0] == 'O' ? 2 : 3), w[2], w[3], w[4]};
            w.swap(v);
        }
        if (digits.empty() || name.empty() || w.size() != 6) continue;
        out.push_back({parse_i(w[2]), w[1], name, parse_i(digits),
                       {parse_f_via_double(w[3]), parse_f_via_double(w[4]), parse_f_via_double(w[5])}});
    }
    return out;
}

struct Where {
    int compound = -1, local = -1;
};

struct CompoundBuild {
    std::vector<F3> pos;
    std::vector<uint8_t> types;
    std::vector<int> serial;
    std::vector<lima_pair_bond> bonds;
    std::vector<lima_angle_bond> angles;
    std::vector<lima_dihedral_bond> dihedrals;
    std::vector<std::vector<std::pair<int, int>>> exclusions;   // per particle: (local, compound)
};

struct BridgeBuild {
    int left, right;
    std::vector<std::pair<int, Where>> particles;                // serial, place
    std::vector<lima_pair_bond> bonds;
    std::vector<lima_angle_bond> angles;
    std::vector<lima_dihedral_bond> dihedrals;
};

// glibc's rand() is random_r on a 128-byte additive-feedback state; a private state of
// the same kind, seeded like srand(seed), yields the same sequence without touching the
// process-wide generator.
class CRand {
public:
    explicit CRand(unsigned seed) {
        std::memset(&data_, 0, sizeof data_);
        if (initstate_r(seed, state_, sizeof state_, &data_) != 0) bad("initstate_r failed");
    }
    int next() {
        int32_t v = 0;
        random_r(&data_, &v);
        draws_++;
        return v;
    }
    int64_t draws() const { return draws_; }

private:
    char state_[128];
    random_data data_;
    int64_t draws_ = 0;
};

// BoxBuilder::random (BoxBuilder.cuh:100-102): rand() % 10000 mapped onto [-1, 1) in FP32.
float rand_to_unit(int r) { return r % 10000 / 10000.f * 2 - 1.f; }

// Three of those, z first (the order GCC and MSVC evaluate Float3(random(), random(),
// random())), normalised (QuantomTypes.cuh:50-55).
F3 random_direction(CRand& rng) {
    auto unit = [&] { return rand_to_unit(rng.next()); };
    float z = unit(), y = unit(), x = unit();
    F3 v{x, y, z};
    float l = length(v);
    if (l == 0.f) return {};
    return scale(v, 1.f / l);
}

// Points binned on a 0.5 nm grid (coarser than both placement radii), so that the
// "nothing within r" tests visit only neighbouring bins.
class PointBins {
public:
    void add(F3 p) {
        pts_.push_back(p);
        bins_[key(cell(p.x), cell(p.y), cell(p.z))].push_back((int)pts_.size() - 1);
    }
    // True when some stored point q has length(q - p) < r (FP32 length, FP64 compare).
    bool any_within(F3 p, double r) const {
        const long cx = cell(p.x), cy = cell(p.y), cz = cell(p.z);
        for (long dz = -1; dz <= 1; dz++)
            for (long dy = -1; dy <= 1; dy++)
                for (long dx = -1; dx <= 1; dx++) {
                    auto it = bins_.find(key(cx + dx, cy + dy, cz + dz));
                    if (it == bins_.end()) continue;
                    for (int i : it->second)
                        if ((double)length(pts_[i] - p) < r) return true;
                }
        return false;
    }

private:
    static long cell(float v) { return (long)std::floor((double)v / kBin); }
    static uint64_t key(long x, long y, long z) {
        return ((uint64_t)(x & 0x1FFFFF) << 42) | ((uint64_t)(y & 0x1FFFFF) << 21) | (uint64_t)(z & 0x1FFFFF);
    }
    static constexpr double kBin = 0.5;
    std::vector<F3> pts_;
    std::unordered_map<uint64_t, std::vector<int>> bins_;
};

}  // namespace

// ------------------------------------------------------------------ the handle

struct lima_ingest {
    lima_ingest_params params{};
    lima_forcefield forcefield{};
    int n_types = 0;
    int n_candidates = 0;
    int64_t draws = 0;
    F3 offset;
    std::vector<int32_t> compound_offset, key, bond_offset, angle_offset, dihedral_offset;
    std::vector<float> sphere, pos, prev;
    std::vector<uint8_t> types;
    std::vector<int32_t> serial;
    std::vector<lima_pair_bond> bonds;
    std::vector<lima_angle_bond> angles;
    std::vector<lima_dihedral_bond> dihedrals;
    std::vector<int32_t> ignore_offset, ignore_compound, ignore_local;
    std::vector<int32_t> bp_offset, bp_compound, bp_local, bb_offset, ba_offset, bd_offset;
    std::vector<lima_pair_bond> bbonds;
    std::vector<lima_angle_bond> bangles;
    std::vector<lima_dihedral_bond> bdihedrals;
    std::vector<float> solvent, solvent_prev;
};

namespace {

void build(lima_ingest& out, const std::string& conf, const std::string& top, const std::string& ff_dir) {
    const lima_ingest_params& P = out.params;
    EngineFF ff = read_filtered(ff_dir + "/LIMA_ffnonbonded_filtered.txt", ff_dir + "/LIMA_ffbonded_filtered.txt");
    out.forcefield = ff.params;
    out.n_types = ff.n_types;
    const std::vector<GroAtom> atoms = read_gro(conf);

    // Residues fill a compound until the next residue might not fit; each new compound is
    // bridged to the one before (CompoundBuilder::loadParticles, CompoundBuilder.cpp:83-150).
    std::vector<CompoundBuild> comps;
    std::vector<BridgeBuild> bridges;
    std::unordered_map<int, Where> where;
    int residue = -1;
    for (const GroAtom& a : atoms) {
        if (a.residue_id < P.min_residue_id) continue;
        if (a.residue_id > P.max_residue_id) break;
        if (P.ignore_hydrogens && a.name[0] == 'H') continue;
        if (a.residue == "SOL") continue;
        if (a.residue_id != residue) {
            if (comps.empty() || (int)comps.back().pos.size() + P.max_atoms_in_residue > P.max_compound_particles) {
                if (!comps.empty()) bridges.push_back({(int)comps.size() - 1, (int)comps.size(), {}, {}, {}, {}});
                comps.emplace_back();
            }
            residue = a.residue_id;
        }
        CompoundBuild& c = comps.back();
        if ((int)c.pos.size() == P.max_compound_particles) throw IngestError(LIMA_E_CAPACITY, "compound full");
        if (a.serial <= 0 || a.serial >= (int)ff.type_of.size() || ff.type_of[a.serial] < 0)
            bad("atom " + std::to_string(a.serial) + " has no type in the filtered force field");
        int t = ff.type_of[a.serial];
        if (t < 0 || t >= LIMA_MAX_ATOM_TYPES) bad("atom type out of range");
        where[a.serial] = {(int)comps.size() - 1, (int)c.pos.size()};
        c.pos.push_back(a.pos);
        c.types.push_back((uint8_t)t);
        c.serial.push_back(a.serial);
        c.exclusions.emplace_back();
    }
    if (comps.empty()) bad("no molecule atoms selected from " + conf);
    if (comps.size() > 255) throw IngestError(LIMA_E_CAPACITY, "more than 255 compounds (u8 exclusion ids)");

    // Bonded terms in topology order (CompoundBuilder::loadTopology / addBond / addAngle /
    // addDihedral, CompoundBuilder.cpp:152-318): a term with an absent atom is dropped;
    // every atom of a kept term excludes every other; a term inside one compound goes to
    // that compound with compound-local indices, a term across two goes to their bridge
    // with bridge-local indices.  Only the first [ dihedrals ] section is read.
    std::vector<std::string> lines = file_lines(top);
    int mode = 0, dihedral_sections = 0;
    for (const std::string& line : lines) {
        Row rec = gro_words(line);
        if (!rec.empty() && rec[0] == ";") continue;
        if (rec.empty()) { mode = 0; continue; }
        if (mode == 0) {
            mode = rec[0] == "bonds" ? 2 : rec[0] == "angles" ? 3 : rec[0] == "dihedrals" ? 4 : 0;
            if (mode == 4) dihedral_sections++;
            continue;
        }
        if (mode == 4 && dihedral_sections > 1) continue;
        if ((int)rec.size() < mode) bad("short topology row: " + line);
        std::array<int, 4> ids{};
        std::array<Where, 4> at{};
        bool present = true;
        for (int k = 0; k < mode; k++) {
            ids[k] = parse_i(rec[k]);
            auto it = where.find(ids[k]);
            if (it == where.end()) { present = false; break; }
            at[k] = it->second;
        }
        if (!present) continue;
        float pa = 0.f, pb = 0.f;
        int pn = 0;
        if (mode == 2) {
            auto it = ff.bonds.find({ids[0], ids[1]});
            if (it == ff.bonds.end()) bad("bond " + std::to_string(ids[0]) + "-" + std::to_string(ids[1]) + " has no parameters");
            pa = it->second.first; pb = it->second.second;
        } else if (mode == 3) {
            auto it = ff.angles.find({ids[0], ids[1], ids[2]});
            if (it == ff.angles.end()) bad("angle has no parameters: " + line);
            pa = it->second.first; pb = it->second.second;
        } else {
            auto it = ff.dihedrals.find({ids[0], ids[1], ids[2], ids[3]});
            if (it == ff.dihedrals.end()) bad("dihedral has no parameters: " + line);
            pa = std::get<0>(it->second); pb = std::get<1>(it->second); pn = std::get<2>(it->second);
        }
        for (int s = 0; s < mode; s++)                       // LJ_Ignores::addIgnoreTarget (Bodies.cuh:176-189)
            for (int o = 0; o < mode; o++) {
                if (s == o) continue;
                auto& ex = comps[at[s].compound].exclusions[at[s].local];
                std::pair<int, int> tgt{at[o].local, at[o].compound};
                if (std::find(ex.begin(), ex.end(), tgt) != ex.end()) continue;
                if ((int)ex.size() == P.max_exclusions) throw IngestError(LIMA_E_CAPACITY, "exclusion list full");
                ex.push_back(tgt);
            }
        // The compounds a term spans: the first atom's, and the last different one seen
        // (GenericBond, Bodies.cuh:118-131).
        int c0 = at[0].compound, c1 = -1;
        for (int k = 1; k < mode; k++)
            if (at[k].compound != c0) c1 = at[k].compound;
        std::array<uint32_t, 4> loc{};
        if (c1 < 0) {
            CompoundBuild& c = comps[c0];
            for (int k = 0; k < mode; k++) loc[k] = (uint32_t)at[k].local;
            if (mode == 2) {
                if (c.bonds.size() == 128) throw IngestError(LIMA_E_CAPACITY, "more than 128 bonds in a compound");
                c.bonds.push_back({pa, pb, {loc[0], loc[1]}});
            } else if (mode == 3) {
                if (c.angles.size() == 256) throw IngestError(LIMA_E_CAPACITY, "more than 256 angles in a compound");
                c.angles.push_back({pa, pb, {loc[0], loc[1], loc[2]}});
            } else {
                if (c.dihedrals.size() == 384) throw IngestError(LIMA_E_CAPACITY, "more than 384 dihedrals in a compound");
                c.dihedrals.push_back({pa, pb, (uint32_t)pn, {loc[0], loc[1], loc[2], loc[3]}});
            }
            continue;
        }
        const int lo = std::min(c0, c1), hi = std::max(c0, c1);
        auto br = std::find_if(bridges.begin(), bridges.end(), [&](const BridgeBuild& b) { return b.left == lo && b.right == hi; });
        if (br == bridges.end())
            bad("a term joins compounds " + std::to_string(lo) + " and " + std::to_string(hi) + ", which share no bridge");
        for (int k = 0; k < mode; k++) {                      // CompoundBridge::addBondParticles (Bodies.cuh:538-557)
            auto it = std::find_if(br->particles.begin(), br->particles.end(), [&](const auto& p) { return p.first == ids[k]; });
            if (it == br->particles.end()) {
                if ((int)br->particles.size() == P.max_bridge_particles) throw IngestError(LIMA_E_CAPACITY, "bridge full");
                br->particles.push_back({ids[k], at[k]});
                it = br->particles.end() - 1;
            }
            loc[k] = (uint32_t)(it - br->particles.begin());
        }
        if (mode == 2) {
            if (br->bonds.size() == 32) throw IngestError(LIMA_E_CAPACITY, "more than 32 bonds in a bridge");
            br->bonds.push_back({pa, pb, {loc[0], loc[1]}});
        } else if (mode == 3) {
            if (br->angles.size() == 32) throw IngestError(LIMA_E_CAPACITY, "more than 32 angles in a bridge");
            br->angles.push_back({pa, pb, {loc[0], loc[1], loc[2]}});
        } else {
            if (br->dihedrals.size() == 32) throw IngestError(LIMA_E_CAPACITY, "more than 32 dihedrals in a bridge");
            br->dihedrals.push_back({pa, pb, (uint32_t)pn, {loc[0], loc[1], loc[2], loc[3]}});
        }
    }

    // Key particle and confining sphere on the file's coordinates (Compound::calcParticleSphere,
    // Bodies.cuh:407-423): the particle nearest the centre (first on ties), and the largest distance.
    for (const CompoundBuild& c : comps) {
        F3 com = centre(c.pos);
        float nearest = INFINITY, furthest = -INFINITY;
        int key = 0;
        for (size_t i = 0; i < c.pos.size(); i++) {
            float d = length(c.pos[i] - com);
            if (d < nearest) key = (int)i;
            nearest = std::min(nearest, d);
            furthest = std::max(furthest, d);
        }
        out.key.push_back(key);
        out.sphere.push_back(furthest);
    }

    // Centre the molecule (BoxBuilder::addSingleMolecule, BoxBuilder.cu:37-63; Molecule::calcCOM,
    // Bodies.cuh:485-491), compounds at rest; BoxBuilder still draws a direction per compound
    // (integrateCompound, BoxBuilder.cu:248-269).
    CRand rng(P.seed);
    F3 mcom;
    const float wc = 1.f / (float)comps.size();
    for (const CompoundBuild& c : comps) mcom = mcom + scale(centre(c.pos), wc);
    const float half = P.box_len / 2.f;
    out.offset = F3{half, half, half} - mcom;
    PointBins chain_bins;
    for (CompoundBuild& c : comps) {
        for (F3& p : c.pos) {
            p = p + out.offset;
            chain_bins.add(p);
        }
        random_direction(rng);
    }

    // The .gro waters' oxygens, shifted like the molecule, kept where no chain particle is
    // within 0.2 nm and no kept solvent within 0.25 nm, each with velocity v_rms in a random
    // direction (BoxBuilder.cu:194-218, 273-276, 440-461; CompoundBuilder.cpp:51-62).
    const double solvent_mass = 18.01528f * 1e-3;                    // Simulation.cuh:30
    const double v_rms = std::sqrt(3 * 8.3144 * 313 / solvent_mass);  // BoxBuilder.cuh:63-71
    PointBins solvent_bins;
    if (P.add_solvents) {
        for (const GroAtom& a : atoms) {
            if (a.residue != "SOL" || a.name[0] != 'O') continue;
            out.n_candidates++;
            if (out.solvent.size() / 3 == 0xFFFF) throw IngestError(LIMA_E_CAPACITY, "more than 65535 solvents");
            F3 p = a.pos + out.offset;
            if (chain_bins.any_within(p, 0.2) || solvent_bins.any_within(p, 0.25)) continue;
            F3 v = scale(random_direction(rng), v_rms);
            F3 q = p - scale(v, P.dt);
            solvent_bins.add(p);
            out.solvent.insert(out.solvent.end(), {p.x, p.y, p.z});
            out.solvent_prev.insert(out.solvent_prev.end(), {q.x, q.y, q.z});
        }
    }
    out.draws = rng.draws();

    // Flatten into the lima_system arrays.
    auto push_offset = [](std::vector<int32_t>& v, size_t n) {
        if (v.empty()) v.push_back(0);
        v.push_back(v.back() + (int32_t)n);
    };
    for (size_t ci = 0; ci < comps.size(); ci++) {
        const CompoundBuild& c = comps[ci];
        push_offset(out.compound_offset, c.pos.size());
        push_offset(out.bond_offset, c.bonds.size());
        push_offset(out.angle_offset, c.angles.size());
        push_offset(out.dihedral_offset, c.dihedrals.size());
        for (size_t i = 0; i < c.pos.size(); i++) {
            out.pos.insert(out.pos.end(), {c.pos[i].x, c.pos[i].y, c.pos[i].z});
            out.types.push_back(c.types[i]);
            out.serial.push_back(c.serial[i]);
            push_offset(out.ignore_offset, c.exclusions[i].size());
            for (const auto& e : c.exclusions[i]) {
                out.ignore_local.push_back(e.first);
                out.ignore_compound.push_back(e.second);
            }
        }
        out.bonds.insert(out.bonds.end(), c.bonds.begin(), c.bonds.end());
        out.angles.insert(out.angles.end(), c.angles.begin(), c.angles.end());
        out.dihedrals.insert(out.dihedrals.end(), c.dihedrals.begin(), c.dihedrals.end());
    }
    out.prev = out.pos;
    out.bp_offset = {0};
    out.bb_offset = {0};
    out.ba_offset = {0};
    out.bd_offset = {0};
    for (const BridgeBuild& b : bridges) {
        push_offset(out.bp_offset, b.particles.size());
        push_offset(out.bb_offset, b.bonds.size());
        push_offset(out.ba_offset, b.angles.size());
        push_offset(out.bd_offset, b.dihedrals.size());
        for (const auto& p : b.particles) {
            out.bp_compound.push_back(p.second.compound);
            out.bp_local.push_back(p.second.local);
        }
        out.bbonds.insert(out.bbonds.end(), b.bonds.begin(), b.bonds.end());
        out.bangles.insert(out.bangles.end(), b.angles.begin(), b.angles.end());
        out.bdihedrals.insert(out.bdihedrals.end(), b.dihedrals.begin(), b.dihedrals.end());
    }
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

const char* lima_ingest_last_error(void) { return g_last_error.c_str(); }

int lima_ff_merge(const char* const* prm_paths, int32_t n_paths, const char* out_dir) {
    try {
        if (!prm_paths || n_paths <= 0 || !out_dir) bad("lima_ff_merge: no input files or no output directory");
        std::vector<std::string> paths;
        for (int i = 0; i < n_paths; i++) {
            if (!prm_paths[i]) bad("lima_ff_merge: null path");
            paths.emplace_back(prm_paths[i]);
        }
        ParamTables t = merge_prm(paths);
        write_text(std::string(out_dir) + "/LIMA_ffnonbonded.txt", merged_nonbonded_text(t));
        write_text(std::string(out_dir) + "/LIMA_ffbonded.txt", merged_bonded_text(t));
        return LIMA_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int lima_ff_prepare(const char* conf_gro, const char* topol_top, const char* ff_dir) {
    try {
        if (!conf_gro || !topol_top || !ff_dir) bad("lima_ff_prepare: null path");
        const std::string dir(ff_dir);
        ParamTables t = read_merged(dir + "/LIMA_ffnonbonded.txt", dir + "/LIMA_ffbonded.txt");
        Filtered f = filter_system(conf_gro, topol_top, t);
        write_text(dir + "/LIMA_ffnonbonded_filtered.txt", filtered_nonbonded_text(f));
        write_text(dir + "/LIMA_ffbonded_filtered.txt", filtered_bonded_text(f));
        return LIMA_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

void lima_ingest_default_params(lima_ingest_params* p) {
    if (!p) return;
    p->min_residue_id = 0;
    p->max_residue_id = 200;
    p->ignore_hydrogens = 1;
    p->max_compound_particles = 64;
    p->max_atoms_in_residue = 32;
    p->max_bridge_particles = 32;
    p->max_exclusions = 32;
    p->add_solvents = 1;
    p->box_len = 7.f;
    p->dt = 0.1 * 1e-6;
    p->thermostat_target = 310.f;
    p->seed = 290128309u;
}

int lima_ingest_create(const char* conf_gro, const char* topol_top, const char* ff_dir,
                       const lima_ingest_params* params, lima_ingest** out) {
    try {
        if (!out) bad("lima_ingest_create: null output");
        *out = nullptr;
        if (!conf_gro || !topol_top || !ff_dir) bad("lima_ingest_create: null path");
        auto h = std::make_unique<lima_ingest>();
        if (params) h->params = *params;
        else lima_ingest_default_params(&h->params);
        const lima_ingest_params& P = h->params;
        if (P.max_compound_particles < 1 || P.max_compound_particles > 128)
            bad("max_compound_particles must be in [1, 128]");
        if (P.max_atoms_in_residue < 0 || P.max_bridge_particles < 1 || P.max_exclusions < 0)
            bad("capacities must be positive");
        if (!(P.box_len > 0.f) || !(P.dt > 0.0)) bad("box_len and dt must be > 0");
        build(*h, conf_gro, topol_top, ff_dir);
        *out = h.release();
        return LIMA_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int lima_ingest_system(const lima_ingest* h, lima_system* s) {
    if (!h || !s) { g_last_error = "lima_ingest_system: null argument"; return LIMA_E_ARG; }
    std::memset(s, 0, sizeof *s);
    s->box_len = h->params.box_len;
    s->dt = h->params.dt;
    s->thermostat_target = h->params.thermostat_target;
    s->forcefield = h->forcefield;
    s->n_compounds = (int32_t)h->key.size();
    s->compound_offset = h->compound_offset.data();
    s->key_particle_index = h->key.data();
    s->confining_particle_sphere = h->sphere.data();
    s->chain_positions = h->pos.data();
    s->chain_prev_positions = h->prev.data();
    s->chain_atom_types = h->types.data();
    s->bond_offset = h->bond_offset.data();
    s->bonds = h->bonds.data();
    s->angle_offset = h->angle_offset.data();
    s->angles = h->angles.data();
    s->dihedral_offset = h->dihedral_offset.data();
    s->dihedrals = h->dihedrals.data();
    s->ignore_offset = h->ignore_offset.data();
    s->ignore_compound = h->ignore_compound.data();
    s->ignore_local = h->ignore_local.data();
    s->n_bridges = (int32_t)h->bp_offset.size() - 1;
    s->bridge_particle_offset = h->bp_offset.data();
    s->bridge_particle_compound = h->bp_compound.data();
    s->bridge_particle_local = h->bp_local.data();
    s->bridge_bond_offset = h->bb_offset.data();
    s->bridge_bonds = h->bbonds.data();
    s->bridge_angle_offset = h->ba_offset.data();
    s->bridge_angles = h->bangles.data();
    s->bridge_dihedral_offset = h->bd_offset.data();
    s->bridge_dihedrals = h->bdihedrals.data();
    s->n_solvents = (int32_t)(h->solvent.size() / 3);
    s->solvent_positions = h->solvent.data();
    s->solvent_prev_positions = h->solvent_prev.data();
    return LIMA_OK;
}

int lima_ingest_get_info(const lima_ingest* h, lima_ingest_info* o) {
    if (!h || !o) { g_last_error = "lima_ingest_get_info: null argument"; return LIMA_E_ARG; }
    o->n_compounds = (int32_t)h->key.size();
    o->n_bridges = (int32_t)h->bp_offset.size() - 1;
    o->n_chain = (int32_t)h->types.size();
    o->n_solvents = (int32_t)(h->solvent.size() / 3);
    o->n_atom_types = h->n_types;
    o->n_solvent_candidates = h->n_candidates;
    o->n_random_draws = h->draws;
    o->offset[0] = h->offset.x;
    o->offset[1] = h->offset.y;
    o->offset[2] = h->offset.z;
    return LIMA_OK;
}

int lima_ingest_global_ids(const lima_ingest* h, int32_t* out) {
    if (!h || !out) { g_last_error = "lima_ingest_global_ids: null argument"; return LIMA_E_ARG; }
    std::copy(h->serial.begin(), h->serial.end(), out);
    return LIMA_OK;
}

void lima_ingest_destroy(lima_ingest* h) { delete h; }

}  // extern "C"
