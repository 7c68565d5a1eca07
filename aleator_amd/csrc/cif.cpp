// cif.cpp -- CIF structure reader with symmetry expansion, the host-side input stage of the
// path: a framework read here is the one the reference's io::readCif returns (same atom order,
// same Cartesian positions bit for bit, same species numbering), so that device runs started
// from a CIF follow the reference's trajectories.
//
// Reference interfaces and the behaviour followed:
//   CIF text  -> document   io/cif_document.cc:185-350 (parseCifDocument), line tokens :76-110,
//                           numbers :160-183 (tryParseCifNumber)
//   "x,y,z"   -> affine op  io/symmetry_operation.cc:33-150
//   labels    -> elements   io/element_table.cc:20-101 (elementFromLabel, atomicMass)
//   document  -> structure  io/structure_io.cc:32-250 (buildLattice, symmetry loop choice,
//                           orbit expansion with a 1e-3 A duplicate test, species by first
//                           appearance of each element)
// Malformed input raises ALEATOR_PARSE_ERROR (io::CifParseError, io/cif_error.hpp:8-16).
#include <algorithm>
#include <array>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numbers>
#include <optional>
#include <sstream>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

#include "aleator_gpu.h"
#include "internal.hpp"

struct aleator_cif_document;
struct aleator_structure;

namespace aleator_gpu::cif {
namespace {

[[noreturn]] void parse_error(const std::string& msg) { fail(ALEATOR_PARSE_ERROR, msg); }
[[noreturn]] void parse_error(const std::string& msg, size_t line) {
    fail(ALEATOR_PARSE_ERROR, msg + " (line " + std::to_string(line) + ")");
}

// Field separators of a CIF line.  A line counts as blank when every byte is isspace().
bool sep(char c) { return c == ' ' || c == '\t' || c == '\r'; }
bool blank(const std::string& s) {
    for (char c : s)
        if (!std::isspace((unsigned char)c)) return false;
    return true;
}
size_t first_non_sep(const std::string& s) {
    size_t k = 0;
    while (k < s.size() && sep(s[k])) ++k;
    return k;
}
std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}
bool has_prefix_nocase(std::string_view s, std::string_view p) {
    if (s.size() < p.size()) return false;
    for (size_t k = 0; k < p.size(); ++k)
        if (std::tolower((unsigned char)s[k]) != std::tolower((unsigned char)p[k])) return false;
    return true;
}

// One line's fields: separator-delimited, '#' outside a field starts a comment, and a quoted
// field ends only at its quote character when that is followed by a separator or the line end.
std::vector<std::string> fields(const std::string& line, size_t lineno) {
    std::vector<std::string> out;
    size_t i = 0;
    const size_t n = line.size();
    for (;;) {
        while (i < n && sep(line[i])) ++i;
        if (i == n || line[i] == '#') return out;
        const char q = line[i];
        if (q == '\'' || q == '"') {
            size_t j = i + 1;
            while (j < n && !(line[j] == q && (j + 1 == n || sep(line[j + 1])))) ++j;
            if (j == n) parse_error("unterminated quoted string", lineno);
            out.emplace_back(line, i + 1, j - i - 1);
            i = j + 1;
        } else {
            size_t j = i;
            while (j < n && !sep(line[j]) && line[j] != '#') ++j;
            out.emplace_back(line, i, j - i);
            i = j;
        }
    }
}

}  // namespace

struct Loop {
    std::vector<std::string> columns;  // lowercased tags
    std::vector<std::vector<std::string>> rows;
    int column(std::string_view tag) const {
        for (size_t c = 0; c < columns.size(); ++c)
            if (columns[c] == tag) return (int)c;
        return -1;
    }
};

struct Document {
    std::string block;
    std::unordered_map<std::string, std::string> tags;  // lowercased tag -> raw value
    std::vector<Loop> loops;

    const std::string* tag(const std::string& name) const {
        auto it = tags.find(lower(name));
        return it == tags.end() ? nullptr : &it->second;
    }
    // First loop that holds every one of `need`.
    const Loop* loop_with(std::initializer_list<std::string_view> need) const {
        for (const Loop& l : loops)
            if (std::all_of(need.begin(), need.end(), [&](std::string_view t) { return l.column(t) >= 0; }))
                return &l;
        return nullptr;
    }
};

namespace {

// Line-oriented reader of the first data block.
class Parser {
public:
    explicit Parser(std::string_view text) {
        size_t start = 0;
        for (;;) {
            const size_t nl = text.find('\n', start);
            std::string line(text.substr(start, nl == std::string_view::npos ? std::string_view::npos : nl - start));
            if (!line.empty() && line.back() == '\r') line.pop_back();
            lines_.push_back(std::move(line));
            if (nl == std::string_view::npos) break;
            start = nl + 1;
        }
    }

    Document run() {
        bool in_block = false;
        while (at_ < lines_.size()) {
            const std::string& line = lines_[at_];
            const size_t lineno = at_ + 1;
            if (blank(line)) {
                ++at_;
                continue;
            }
            const std::string_view body = std::string_view(line).substr(first_non_sep(line));
            if (body[0] == '#') {
                ++at_;
            } else if (has_prefix_nocase(body, "data_")) {
                if (in_block) break;  // only the first data block is read
                const auto f = fields(line, lineno);
                doc_.block = f.empty() ? std::string() : f[0].substr(5);
                in_block = true;
                ++at_;
            } else if (has_prefix_nocase(body, "loop_")) {
                read_loop(lineno);
            } else if (body[0] == '_') {
                read_tag(line, lineno);
            } else {
                parse_error("unrecognized content outside of a tag, loop_, or data_ block", lineno);
            }
        }
        if (!in_block) parse_error("no data_ block found in file");
        return std::move(doc_);
    }

private:
    // A line that starts a tag, a loop_ or a data_ block (comments and blanks do not).
    static bool starts_construct(const std::string& line) {
        const size_t k = first_non_sep(line);
        if (k == line.size() || line[k] == '#') return false;
        const std::string_view r = std::string_view(line).substr(k);
        return r[0] == '_' || has_prefix_nocase(r, "loop_") || has_prefix_nocase(r, "data_");
    }

    void read_loop(size_t loop_line) {
        ++at_;
        Loop loop;
        // column tags: the run of lines whose first field starts with '_'
        for (; at_ < lines_.size(); ++at_) {
            const std::string& line = lines_[at_];
            if (blank(line)) continue;
            const size_t k = first_non_sep(line);
            if (k < line.size() && line[k] == '#') continue;
            if (k == line.size() || line[k] != '_') break;
            const auto f = fields(line, at_ + 1);
            if (f.empty()) parse_error("empty loop_ column declaration", at_ + 1);
            loop.columns.push_back(lower(f[0]));
        }
        if (loop.columns.empty()) parse_error("loop_ with no column declarations", loop_line);
        // values: whitespace-separated across lines, grouped into rows of columns.size()
        std::vector<std::string> row;
        size_t last_line = loop_line;
        for (; at_ < lines_.size(); ++at_) {
            const std::string& line = lines_[at_];
            if (blank(line)) continue;
            if (row.empty() && starts_construct(line)) break;
            last_line = at_ + 1;
            for (auto& v : fields(line, at_ + 1)) {
                row.push_back(std::move(v));
                if (row.size() == loop.columns.size()) {
                    loop.rows.push_back(std::move(row));
                    row.clear();
                }
            }
        }
        if (!row.empty())
            parse_error("loop_ row has " + std::to_string(row.size()) + " value(s) but " +
                            std::to_string(loop.columns.size()) + " columns",
                        last_line);
        doc_.loops.push_back(std::move(loop));
    }

    void read_tag(const std::string& line, size_t lineno) {
        const auto f = fields(line, lineno);
        const std::string tag = lower(f[0]);
        ++at_;
        if (f.size() >= 2) {
            doc_.tags[tag] = f[1];
            return;
        }
        if (at_ == lines_.size()) parse_error("missing value for tag " + tag + " at end of file", lineno);
        const std::string& next = lines_[at_];
        if (!next.empty() && next[0] == ';') {
            // semicolon text field: up to the next line that starts with ';'
            const size_t open = at_ + 1;
            std::string value = next.substr(1);
            bool have = !value.empty();
            for (++at_; at_ < lines_.size(); ++at_) {
                if (!lines_[at_].empty() && lines_[at_][0] == ';') {
                    ++at_;
                    doc_.tags[tag] = std::move(value);
                    return;
                }
                if (have) value += '\n';
                value += lines_[at_];
                have = true;
            }
            parse_error("unterminated ';' text field for tag " + tag, open);
        }
        const auto v = fields(next, at_ + 1);
        if (v.empty()) parse_error("missing value for tag " + tag, lineno);
        doc_.tags[tag] = v[0];
        ++at_;
    }

    std::vector<std::string> lines_;
    size_t at_ = 0;
    Document doc_;
};

// ---------------------------------------------------------------- elements
// Standard atomic weights (amu) of the elements a framework file names; symbols in canonical
// case.  The same 83 entries and values as io/element_table.cc:20-45.
struct Element {
    const char* symbol;
    double mass;
};
constexpr Element kElements[] = {
    {"H", 1.008},     {"He", 4.003},    {"Li", 6.94},     {"Be", 9.012},    {"B", 10.81},
    {"C", 12.011},    {"N", 14.007},    {"O", 15.999},    {"F", 18.998},    {"Ne", 20.180},
    {"Na", 22.990},   {"Mg", 24.305},   {"Al", 26.982},   {"Si", 28.085},   {"P", 30.974},
    {"S", 32.06},     {"Cl", 35.45},    {"Ar", 39.948},   {"K", 39.098},    {"Ca", 40.078},
    {"Sc", 44.956},   {"Ti", 47.867},   {"V", 50.942},    {"Cr", 51.996},   {"Mn", 54.938},
    {"Fe", 55.845},   {"Co", 58.933},   {"Ni", 58.693},   {"Cu", 63.546},   {"Zn", 65.38},
    {"Ga", 69.723},   {"Ge", 72.630},   {"As", 74.922},   {"Se", 78.971},   {"Br", 79.904},
    {"Kr", 83.798},   {"Rb", 85.468},   {"Sr", 87.62},    {"Y", 88.906},    {"Zr", 91.224},
    {"Nb", 92.906},   {"Mo", 95.95},    {"Tc", 98.0},     {"Ru", 101.07},   {"Rh", 102.906},
    {"Pd", 106.42},   {"Ag", 107.868},  {"Cd", 112.414},  {"In", 114.818},  {"Sn", 118.710},
    {"Sb", 121.760},  {"Te", 127.60},   {"I", 126.904},   {"Xe", 131.293},  {"Cs", 132.905},
    {"Ba", 137.327},  {"La", 138.905},  {"Ce", 140.116},  {"Pr", 140.908},  {"Nd", 144.242},
    {"Pm", 145.0},    {"Sm", 150.36},   {"Eu", 151.964},  {"Gd", 157.25},   {"Tb", 158.925},
    {"Dy", 162.500},  {"Ho", 164.930},  {"Er", 167.259},  {"Tm", 168.934},  {"Yb", 173.045},
    {"Lu", 174.967},  {"Hf", 178.49},   {"Ta", 180.948},  {"W", 183.84},    {"Re", 186.207},
    {"Os", 190.23},   {"Ir", 192.217},  {"Pt", 195.084},  {"Au", 196.967},  {"Hg", 200.592},
    {"Tl", 204.38},   {"Pb", 207.2},    {"Bi", 208.980},
};

const Element* element_nocase(std::string_view s) {
    for (const Element& e : kElements)
        if (std::strlen(e.symbol) == s.size() && has_prefix_nocase(e.symbol, s)) return &e;
    return nullptr;
}

}  // namespace

// Leading letters of a label name the element: two letters when they spell one, else one.
std::optional<std::string> element_from_label(std::string_view label) {
    size_t letters = 0;
    while (letters < label.size() && std::isalpha((unsigned char)label[letters])) ++letters;
    if (letters == 0) return std::nullopt;
    if (letters >= 2)
        if (const Element* e = element_nocase(label.substr(0, 2))) return std::string(e->symbol);
    if (const Element* e = element_nocase(label.substr(0, 1))) return std::string(e->symbol);
    return std::nullopt;
}

std::optional<double> atomic_mass(std::string_view symbol) {
    for (const Element& e : kElements)
        if (symbol == e.symbol) return e.mass;
    return std::nullopt;
}

// A CIF number: "?" and "." are absent values, a "(uncertainty)" suffix is dropped, and the
// rest must read completely as a double (strtod syntax, overflow and underflow rejected).
std::optional<double> try_number(const std::string& token) {
    if (token == "?" || token == ".") return std::nullopt;
    const std::string s = token.substr(0, token.find('('));
    if (s.empty()) return std::nullopt;
    errno = 0;
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || errno == ERANGE || end != s.c_str() + s.size()) return std::nullopt;
    return v;
}

double number(const std::string& token) {
    const auto v = try_number(token);
    if (!v) parse_error("expected a number but found \"" + token + "\"");
    return *v;
}

// ---------------------------------------------------------------- symmetry operations
struct SymOp {
    double m[3][3] = {};
    double t[3] = {};
    std::array<double, 3> apply(const std::array<double, 3>& f) const {
        std::array<double, 3> r;
        for (int i = 0; i < 3; ++i) r[i] = m[i][0] * f[0] + m[i][1] * f[1] + m[i][2] * f[2] + t[i];
        return r;
    }
};

// One component such as "1/2-x+y": signed terms, each a number (optionally a fraction), a
// variable, or a number times a variable.
static void parse_component(const std::string& c, const std::string& whole, double row[3], double& shift) {
    if (c.empty()) parse_error("empty component in symmetry operation \"" + whole + "\"");
    size_t i = 0;
    auto is_num = [&](size_t k) { return k < c.size() && (std::isdigit((unsigned char)c[k]) || c[k] == '.'); };
    while (i < c.size()) {
        double sign = 1.0;
        if (c[i] == '+' || c[i] == '-') sign = c[i++] == '-' ? -1.0 : 1.0;
        if (i == c.size()) parse_error("trailing sign with no term in symmetry operation \"" + whole + "\"");
        std::optional<double> value;
        if (is_num(i)) {
            size_t j = i;
            while (is_num(j)) ++j;
            double v = std::strtod(c.substr(i, j - i).c_str(), nullptr);
            i = j;
            if (i < c.size() && c[i] == '/') {
                size_t d = ++i;
                while (d < c.size() && std::isdigit((unsigned char)c[d])) ++d;
                if (d == i) parse_error("malformed fraction in symmetry operation \"" + whole + "\"");
                v = v / std::strtod(c.substr(i, d - i).c_str(), nullptr);
                i = d;
            }
            value = v;
        }
        const char var = i < c.size() ? (char)std::tolower((unsigned char)c[i]) : '\0';
        if (var == 'x' || var == 'y' || var == 'z') {
            row[var - 'x'] += sign * value.value_or(1.0);
            ++i;
        } else if (value) {
            shift += sign * *value;
        } else {
            parse_error("expected x, y, z, or a number in symmetry operation \"" + whole + "\"");
        }
    }
}

SymOp parse_symop(const std::string& text) {
    std::string s;
    for (char ch : text)
        if (!std::isspace((unsigned char)ch)) s += ch;
    std::vector<std::string> parts(1);
    for (char ch : s) {
        if (ch == ',') {
            if (parts.size() == 3)
                parse_error("symmetry operation \"" + text + "\" has more than 3 comma-separated components");
            parts.emplace_back();
        } else {
            parts.back() += ch;
        }
    }
    if (parts.size() != 3)
        parse_error("symmetry operation \"" + text + "\" must have exactly 3 components, found " +
                    std::to_string(parts.size()));
    SymOp op;
    for (int r = 0; r < 3; ++r) parse_component(parts[r], text, op.m[r], op.t[r]);
    return op;
}

// ---------------------------------------------------------------- structure
struct Structure {
    aleator_lattice lattice{};
    std::vector<double> x, y, z, mass, charge, occupancy;
    std::vector<uint32_t> species;
    std::vector<std::string> symbols, labels;
};

namespace {

double cell_value(const Document& d, const char* tag) {
    const std::string* v = d.tag(tag);
    if (!v) parse_error(std::string("missing required tag ") + tag);
    return number(*v);
}

// Row-vector cell from a, b, c (A) and alpha, beta, gamma (degrees): a along x, b in the xy
// plane; the usual crystallographic construction, term order as io/structure_io.cc:58-110.
aleator_lattice cell(const Document& d) {
    const double len[3] = {cell_value(d, "_cell_length_a"), cell_value(d, "_cell_length_b"),
                           cell_value(d, "_cell_length_c")};
    const double deg[3] = {cell_value(d, "_cell_angle_alpha"), cell_value(d, "_cell_angle_beta"),
                           cell_value(d, "_cell_angle_gamma")};
    static const char* const axis[3] = {"a", "b", "c"};
    static const char* const angle[3] = {"alpha", "beta", "gamma"};
    for (int i = 0; i < 3; ++i)
        if (!(len[i] > 0.0))
            parse_error(std::string("_cell_length_") + axis[i] + " must be positive, got " + std::to_string(len[i]));
    for (int i = 0; i < 3; ++i)
        if (!(deg[i] > 0.0 && deg[i] < 180.0))
            parse_error(std::string("_cell_angle_") + angle[i] +
                        " must be strictly between 0 and 180 degrees, got " + std::to_string(deg[i]));
    constexpr double rad = std::numbers::pi / 180.0;
    const double ca = std::cos(deg[0] * rad), cb = std::cos(deg[1] * rad), cg = std::cos(deg[2] * rad);
    const double sg = std::sin(deg[2] * rad);
    const double zz = 1.0 - ca * ca - cb * cb - cg * cg + 2.0 * ca * cb * cg;
    if (zz < 0.0) parse_error("cell angles alpha/beta/gamma do not describe a valid parallelepiped");
    const double c = len[2];
    aleator_lattice m{};
    m.m[0][0] = len[0];
    m.m[1][0] = len[1] * cg;
    m.m[1][1] = len[1] * sg;
    m.m[2][0] = c * cb;
    m.m[2][1] = c * (ca - cb * cg) / sg;
    m.m[2][2] = c * std::sqrt(zz) / sg;
    return m;
}

std::array<double, 3> to_cart(const aleator_lattice& l, const std::array<double, 3>& f) {
    std::array<double, 3> c;
    for (int i = 0; i < 3; ++i) c[i] = f[0] * l.m[0][i] + f[1] * l.m[1][i] + f[2] * l.m[2][i];
    return c;
}

std::vector<SymOp> symmetry(const Document& d) {
    const char* col = "_symmetry_equiv_pos_as_xyz";
    const Loop* l = d.loop_with({col});
    if (!l) {
        col = "_space_group_symop_operation_xyz";
        l = d.loop_with({col});
    }
    if (!l) return {parse_symop("x,y,z")};  // P1
    const int c = l->column(col);
    std::vector<SymOp> ops;
    for (const auto& row : l->rows) ops.push_back(parse_symop(row[c]));
    if (ops.empty()) parse_error(std::string("symmetry loop (") + col + ") contains no operations");
    return ops;
}

}  // namespace

Structure build(const Document& d) {
    Structure s;
    s.lattice = cell(d);
    const std::vector<SymOp> ops = symmetry(d);
    const Loop* atoms = d.loop_with({"_atom_site_label", "_atom_site_fract_x", "_atom_site_fract_y", "_atom_site_fract_z"});
    if (!atoms) parse_error("no _atom_site loop with a label and fractional x/y/z coordinates was found");
    const int c_label = atoms->column("_atom_site_label"), c_type = atoms->column("_atom_site_type_symbol");
    const int c_x = atoms->column("_atom_site_fract_x"), c_y = atoms->column("_atom_site_fract_y"),
              c_z = atoms->column("_atom_site_fract_z");
    const int c_occ = atoms->column("_atom_site_occupancy"), c_q = atoms->column("_atom_site_charge");
    std::unordered_map<std::string, uint32_t> ids;
    for (const auto& row : atoms->rows) {
        const std::string& label = row[c_label];
        std::optional<std::string> el;
        if (c_type >= 0) el = element_from_label(row[c_type]);
        if (!el) el = element_from_label(label);
        if (!el) parse_error("could not determine a chemical element for atom site \"" + label + "\"");
        const auto mass = atomic_mass(*el);
        if (!mass) parse_error("unknown element \"" + *el + "\" (no atomic mass data) for atom site \"" + label + "\"");
        const std::array<double, 3> f0 = {number(row[c_x]), number(row[c_y]), number(row[c_z])};
        const double occ = c_occ >= 0 ? try_number(row[c_occ]).value_or(1.0) : 1.0;
        const double q = c_q >= 0 ? try_number(row[c_q]).value_or(0.0) : 0.0;
        auto [it, fresh] = ids.try_emplace(*el, (uint32_t)s.symbols.size());
        if (fresh) s.symbols.push_back(*el);

        // orbit of the site: images wrapped into [0,1), dropping any within 1e-3 A (through the
        // cell, nearest periodic copy) of one already kept
        std::vector<std::array<double, 3>> orbit;
        for (const SymOp& op : ops) {
            std::array<double, 3> f = op.apply(f0);
            for (double& v : f) v -= std::floor(v);
            bool dup = false;
            for (const auto& o : orbit) {
                std::array<double, 3> df = {f[0] - o[0], f[1] - o[1], f[2] - o[2]};
                for (double& v : df) v -= std::round(v);
                const auto dc = to_cart(s.lattice, df);
                if (std::sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]) < 1e-3) {
                    dup = true;
                    break;
                }
            }
            if (!dup) orbit.push_back(f);
        }
        for (const auto& f : orbit) {
            const auto c = to_cart(s.lattice, f);
            s.x.push_back(c[0]);
            s.y.push_back(c[1]);
            s.z.push_back(c[2]);
            s.mass.push_back(*mass);
            s.charge.push_back(q);
            s.species.push_back(it->second);
            s.occupancy.push_back(occ);
            s.labels.push_back(label);
        }
    }
    return s;
}

Document parse(std::string_view text) { return Parser(text).run(); }

std::string read_file(const char* path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) parse_error(std::string("could not open file: ") + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

}  // namespace aleator_gpu::cif

// ---------------------------------------------------------------- C ABI
using namespace aleator_gpu;

struct aleator_cif_document {
    cif::Document doc;
};
struct aleator_structure {
    cif::Structure s;
};

extern "C" {

aleator_status aleator_cif_document_parse(const char* text, uint64_t length, aleator_cif_document** out) {
    return guarded([&] {
        require(out && (text || length == 0), ALEATOR_INVALID_ARGUMENT, "null argument");
        auto* d = new aleator_cif_document{cif::parse(std::string_view(text ? text : "", length))};
        *out = d;
    });
}

aleator_status aleator_cif_document_destroy(aleator_cif_document* doc) {
    delete doc;
    return ALEATOR_OK;
}

aleator_status aleator_cif_document_block_name(const aleator_cif_document* doc, const char** name) {
    return guarded([&] {
        require(doc && name, ALEATOR_INVALID_ARGUMENT, "null argument");
        *name = doc->doc.block.c_str();
    });
}

aleator_status aleator_cif_document_find_tag(const aleator_cif_document* doc, const char* tag, const char** value) {
    return guarded([&] {
        require(doc && tag && value, ALEATOR_INVALID_ARGUMENT, "null argument");
        const std::string* v = doc->doc.tag(tag);
        *value = v ? v->c_str() : nullptr;
    });
}

aleator_status aleator_cif_document_num_tags(const aleator_cif_document* doc, uint64_t* n) {
    return guarded([&] {
        require(doc && n, ALEATOR_INVALID_ARGUMENT, "null argument");
        *n = doc->doc.tags.size();
    });
}

aleator_status aleator_cif_document_tag(const aleator_cif_document* doc, uint64_t index, const char** tag,
                                        const char** value) {
    return guarded([&] {
        require(doc && tag && value, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(index < doc->doc.tags.size(), ALEATOR_OUT_OF_RANGE, "tag index out of range");
        std::vector<const std::pair<const std::string, std::string>*> order;
        for (const auto& kv : doc->doc.tags) order.push_back(&kv);
        std::sort(order.begin(), order.end(), [](auto* a, auto* b) { return a->first < b->first; });
        *tag = order[index]->first.c_str();
        *value = order[index]->second.c_str();
    });
}

aleator_status aleator_cif_document_num_loops(const aleator_cif_document* doc, uint64_t* n) {
    return guarded([&] {
        require(doc && n, ALEATOR_INVALID_ARGUMENT, "null argument");
        *n = doc->doc.loops.size();
    });
}

aleator_status aleator_cif_document_find_loop(const aleator_cif_document* doc, const char* const* tags,
                                              uint32_t n_tags, int64_t* loop) {
    return guarded([&] {
        require(doc && loop && (tags || n_tags == 0), ALEATOR_INVALID_ARGUMENT, "null argument");
        *loop = -1;
        const auto& loops = doc->doc.loops;
        for (size_t l = 0; l < loops.size() && *loop < 0; ++l) {
            bool all = true;
            for (uint32_t t = 0; t < n_tags && all; ++t) all = loops[l].column(tags[t]) >= 0;
            if (all) *loop = (int64_t)l;
        }
    });
}

aleator_status aleator_cif_document_loop_shape(const aleator_cif_document* doc, uint64_t loop, uint64_t* n_columns,
                                               uint64_t* n_rows) {
    return guarded([&] {
        require(doc && n_columns && n_rows, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(loop < doc->doc.loops.size(), ALEATOR_OUT_OF_RANGE, "loop index out of range");
        *n_columns = doc->doc.loops[loop].columns.size();
        *n_rows = doc->doc.loops[loop].rows.size();
    });
}

aleator_status aleator_cif_document_loop_column(const aleator_cif_document* doc, uint64_t loop, uint64_t column,
                                                const char** tag) {
    return guarded([&] {
        require(doc && tag, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(loop < doc->doc.loops.size() && column < doc->doc.loops[loop].columns.size(), ALEATOR_OUT_OF_RANGE,
                "loop or column index out of range");
        *tag = doc->doc.loops[loop].columns[column].c_str();
    });
}

aleator_status aleator_cif_document_loop_value(const aleator_cif_document* doc, uint64_t loop, uint64_t row,
                                               uint64_t column, const char** value) {
    return guarded([&] {
        require(doc && value, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(loop < doc->doc.loops.size(), ALEATOR_OUT_OF_RANGE, "loop index out of range");
        const auto& l = doc->doc.loops[loop];
        require(row < l.rows.size() && column < l.columns.size(), ALEATOR_OUT_OF_RANGE, "row or column out of range");
        *value = l.rows[row][column].c_str();
    });
}

aleator_status aleator_cif_try_parse_number(const char* token, double* value, int32_t* has_value) {
    return guarded([&] {
        require(token && value && has_value, ALEATOR_INVALID_ARGUMENT, "null argument");
        const auto v = cif::try_number(token);
        *has_value = v.has_value();
        *value = v.value_or(0.0);
    });
}

aleator_status aleator_cif_parse_number(const char* token, double* value) {
    return guarded([&] {
        require(token && value, ALEATOR_INVALID_ARGUMENT, "null argument");
        *value = cif::number(token);
    });
}

aleator_status aleator_parse_symmetry_operation(const char* text, double matrix[9], double translation[3]) {
    return guarded([&] {
        require(text && matrix && translation, ALEATOR_INVALID_ARGUMENT, "null argument");
        const cif::SymOp op = cif::parse_symop(text);
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) matrix[3 * i + j] = op.m[i][j];
            translation[i] = op.t[i];
        }
    });
}

aleator_status aleator_element_from_label(const char* label, char symbol[4], int32_t* found) {
    return guarded([&] {
        require(label && symbol && found, ALEATOR_INVALID_ARGUMENT, "null argument");
        const auto e = cif::element_from_label(label);
        *found = e.has_value();
        std::memset(symbol, 0, 4);
        if (e) std::memcpy(symbol, e->c_str(), e->size());
    });
}

aleator_status aleator_atomic_mass(const char* symbol, double* mass, int32_t* found) {
    return guarded([&] {
        require(symbol && mass && found, ALEATOR_INVALID_ARGUMENT, "null argument");
        const auto m = cif::atomic_mass(symbol);
        *found = m.has_value();
        *mass = m.value_or(0.0);
    });
}

aleator_status aleator_read_cif(const char* path, aleator_structure** out) {
    return guarded([&] {
        require(path && out, ALEATOR_INVALID_ARGUMENT, "null argument");
        const std::string text = cif::read_file(path);
        *out = new aleator_structure{cif::build(cif::parse(text))};
    });
}

aleator_status aleator_parse_cif(const char* text, uint64_t length, aleator_structure** out) {
    return guarded([&] {
        require(out && (text || length == 0), ALEATOR_INVALID_ARGUMENT, "null argument");
        *out = new aleator_structure{cif::build(cif::parse(std::string_view(text ? text : "", length)))};
    });
}

aleator_status aleator_structure_destroy(aleator_structure* s) {
    delete s;
    return ALEATOR_OK;
}

aleator_status aleator_structure_size(const aleator_structure* s, uint64_t* n_atoms, uint32_t* n_species) {
    return guarded([&] {
        require(s && n_atoms && n_species, ALEATOR_INVALID_ARGUMENT, "null argument");
        *n_atoms = s->s.x.size();
        *n_species = (uint32_t)s->s.symbols.size();
    });
}

aleator_status aleator_structure_lattice(const aleator_structure* s, aleator_lattice* lattice) {
    return guarded([&] {
        require(s && lattice, ALEATOR_INVALID_ARGUMENT, "null argument");
        *lattice = s->s.lattice;
    });
}

aleator_status aleator_structure_atoms(const aleator_structure* s, double* x, double* y, double* z, double* mass,
                                       double* charge, uint32_t* species, double* occupancy) {
    return guarded([&] {
        require(s != nullptr, ALEATOR_INVALID_ARGUMENT, "null argument");
        const cif::Structure& t = s->s;
        const size_t n = t.x.size();
        auto put = [n](auto* dst, const auto& src) {
            if (dst && n) std::copy(src.begin(), src.end(), dst);
        };
        put(x, t.x);
        put(y, t.y);
        put(z, t.z);
        put(mass, t.mass);
        put(charge, t.charge);
        put(species, t.species);
        put(occupancy, t.occupancy);
    });
}

aleator_status aleator_structure_species_symbol(const aleator_structure* s, uint32_t species, const char** symbol) {
    return guarded([&] {
        require(s && symbol, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(species < s->s.symbols.size(), ALEATOR_OUT_OF_RANGE, "species index out of range");
        *symbol = s->s.symbols[species].c_str();
    });
}

aleator_status aleator_structure_source_label(const aleator_structure* s, uint64_t atom, const char** label) {
    return guarded([&] {
        require(s && label, ALEATOR_INVALID_ARGUMENT, "null argument");
        require(atom < s->s.labels.size(), ALEATOR_OUT_OF_RANGE, "atom index out of range");
        *label = s->s.labels[atom].c_str();
    });
}

}  // extern "C"
