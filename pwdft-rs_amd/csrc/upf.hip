// UPF v2 pseudopotential reader (host): UpfPseudoPotential::load_from_path
// (src/pseudopotential/upf/mod.rs:98-109) with parse_upf_body
// (src/pseudopotential/upf/parse.rs:20-140) and the block helpers of
// src/pseudopotential/upf/xml.rs. Values come out in the units pwk_species
// documents: r, rab in Å; v_local, D_ij in eV; χ = r·β in Å^-1/2; 4πr²ρ_atom in
// e/Å; ρ_core in e/Å³.
//
// The lookups are plain first-occurrence text searches, as in the reference: an
// attribute is the text between `name="` and the next quote, a data block the
// whitespace-separated numbers between the end of `<TAG ...>` and `</TAG>`, and
// every block must hold exactly the expected count.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

struct pwk_upf {
  std::string element;
  int atomic_number = 0;
  double z_valence = 0.0;
  int l_max = 0;
  int mesh = 0;
  std::vector<double> r, rab, v_local, beta, dij, rho_atom, rho_core;
  std::vector<int32_t> proj_l;
  bool nlcc = false;
};

namespace pwk {
namespace {

constexpr double kRyToEv = 13.605693122994;        // src/consts.rs:17
constexpr double kBohrToAng = 0.529177210903;      // src/consts.rs:19
constexpr double kBohr3ToAng3 = kBohrToAng * kBohrToAng * kBohrToAng;  // src/consts.rs:21

// Element symbols by atomic number (the reference maps the header's symbol through
// elements_rs::Element::from_str).
const char* const kSymbols[] = {
    "",   "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl",
    "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br",
    "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",
    "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu",
    "Hf", "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac",
    "Th", "Pa", "U",  "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr", "Rf", "Db", "Sg", "Bh",
    "Hs", "Mt", "Ds", "Rg", "Cn", "Nh", "Fl", "Mc", "Lv", "Ts", "Og"};

[[noreturn]] void parse_error(const std::string& m) { throw Error(PWK_E_INVALID, "UPF: " + m); }

std::string trim(const std::string& s) {
  const char* ws = " \t\r\n\f\v";
  const size_t a = s.find_first_not_of(ws);
  if (a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(ws);
  return s.substr(a, b - a + 1);
}

// xml.rs extract_attr: text between `name="` and the next quote (false if absent)
bool attr(const std::string& content, const std::string& name, std::string& out) {
  const std::string pat = name + "=\"";
  const size_t s = content.find(pat);
  if (s == std::string::npos) return false;
  const size_t b = s + pat.size();
  const size_t e = content.find('"', b);
  if (e == std::string::npos) return false;
  out = content.substr(b, e - b);
  return true;
}

double parse_f64(const std::string& tok, const std::string& what) {
  const std::string t = trim(tok);
  if (t.empty()) parse_error(what + ": empty number");
  errno = 0;
  char* end = nullptr;
  const double v = std::strtod(t.c_str(), &end);
  if (end != t.c_str() + t.size() || errno == ERANGE) parse_error(what + ": cannot parse '" + t + "'");
  return v;
}

long parse_int(const std::string& tok, const std::string& what) {
  const std::string t = trim(tok);
  if (t.empty()) parse_error(what + ": empty integer");
  errno = 0;
  char* end = nullptr;
  const long v = std::strtol(t.c_str(), &end, 10);
  if (end != t.c_str() + t.size() || errno == ERANGE) parse_error(what + ": cannot parse '" + t + "'");
  return v;
}

std::string required_attr(const std::string& content, const std::string& name) {
  std::string v;
  if (!attr(content, name, v)) parse_error("missing " + name);
  return v;
}

// xml.rs extract_data_block: the numbers of <TAG ...> ... </TAG>, exactly `expected` of them.
// Returns false when the opening tag is absent (the optional blocks treat that as "no data").
bool data_block(const std::string& content, const std::string& tag, size_t expected, std::vector<double>& out) {
  const size_t t = content.find("<" + tag);
  if (t == std::string::npos) return false;
  const size_t gt = content.find('>', t);
  if (gt == std::string::npos) parse_error("malformed tag <" + tag + ">");
  const size_t b = gt + 1;
  const size_t e = content.find("</" + tag + ">", b);
  if (e == std::string::npos) parse_error("missing closing tag </" + tag + ">");
  out.clear();
  out.reserve(expected);
  const char* p = content.data() + b;
  const char* q = content.data() + e;
  while (p < q) {
    while (p < q && std::isspace((unsigned char)*p)) ++p;
    if (p >= q) break;
    const char* s = p;
    while (p < q && !std::isspace((unsigned char)*p)) ++p;
    out.push_back(parse_f64(std::string(s, p), "float parse error in " + tag));
  }
  if (out.size() != expected)
    parse_error(tag + ": expected " + std::to_string(expected) + " values, got " + std::to_string(out.size()));
  return true;
}

void required_block(const std::string& content, const std::string& tag, size_t expected, std::vector<double>& out) {
  if (!data_block(content, tag, expected, out)) parse_error("missing tag <" + tag + ">");
}

// xml.rs extract_beta_angular_momentum
int beta_l(const std::string& content, const std::string& tag) {
  const size_t t = content.find("<" + tag);
  if (t == std::string::npos) parse_error("missing tag <" + tag + ">");
  const size_t gt = content.find('>', t);
  if (gt == std::string::npos) parse_error("malformed tag <" + tag + ">");
  std::string v;
  if (!attr(content.substr(t, gt - t), "angular_momentum", v))
    parse_error("missing angular_momentum attribute in " + tag);
  const long l = parse_int(v, "angular_momentum parse error in " + tag);
  if (l < 0) parse_error(tag + ": angular_momentum must be non-negative (got " + std::to_string(l) + ")");
  return (int)l;
}

std::unique_ptr<pwk_upf> parse_body(const std::string& content) {
  if (content.find("<UPF") == std::string::npos && content.find("<PP_HEADER") == std::string::npos)
    parse_error("unrecognized pseudopotential format (only UPF v2 is supported)");
  auto u = std::make_unique<pwk_upf>();
  u->element = trim(required_attr(content, "element"));
  for (int z = 1; z < (int)(sizeof(kSymbols) / sizeof(kSymbols[0])); ++z)
    if (u->element == kSymbols[z]) u->atomic_number = z;
  if (u->atomic_number == 0) parse_error("unknown element '" + u->element + "'");
  u->z_valence = parse_f64(required_attr(content, "z_valence"), "z_valence parse error");
  u->l_max = (int)parse_int(required_attr(content, "l_max"), "l_max parse error");
  const long nproj = parse_int(required_attr(content, "number_of_proj"), "number_of_proj parse error");
  const long mesh = parse_int(required_attr(content, "mesh_size"), "mesh_size parse error");
  if (nproj < 0 || mesh < 1 || mesh > (1 << 24) || nproj > 64) parse_error("mesh_size or number_of_proj out of range");
  u->mesh = (int)mesh;
  const size_t m = (size_t)mesh;
  std::vector<double> v;
  required_block(content, "PP_R", m, v);
  u->r.resize(m);
  for (size_t i = 0; i < m; ++i) u->r[i] = v[i] * kBohrToAng;
  required_block(content, "PP_RAB", m, v);
  u->rab.resize(m);
  for (size_t i = 0; i < m; ++i) u->rab[i] = v[i] * kBohrToAng;
  required_block(content, "PP_LOCAL", m, v);
  u->v_local.resize(m);
  for (size_t i = 0; i < m; ++i) u->v_local[i] = v[i] * kRyToEv;
  const double sq = std::sqrt(kBohrToAng);
  u->beta.resize((size_t)nproj * m);
  for (long i = 1; i <= nproj; ++i) {
    const std::string tag = "PP_BETA." + std::to_string(i);
    u->proj_l.push_back(beta_l(content, tag));
    required_block(content, tag, m, v);
    for (size_t j = 0; j < m; ++j) u->beta[(size_t)(i - 1) * m + j] = v[j] / sq;
  }
  required_block(content, "PP_DIJ", (size_t)(nproj * nproj), v);
  u->dij.resize(v.size());
  for (size_t i = 0; i < v.size(); ++i) u->dij[i] = v[i] * kRyToEv;
  // 4πr²ρ_atom in e/Bohr → e/Å; zeros when the block is absent
  u->rho_atom.assign(m, 0.0);
  if (data_block(content, "PP_RHOATOM", m, v))
    for (size_t i = 0; i < m; ++i) u->rho_atom[i] = v[i] / kBohrToAng;
  // bare ρ_core in e/Bohr³ → e/Å³, only with the core-correction flag (zeros without the block)
  u->nlcc = content.find("core_correction=\"T\"") != std::string::npos ||
            content.find("core_correction=\"t\"") != std::string::npos ||
            content.find("nlcc=.true.") != std::string::npos;
  if (u->nlcc) {
    u->rho_core.assign(m, 0.0);
    if (data_block(content, "PP_NLCC", m, v))
      for (size_t i = 0; i < m; ++i) u->rho_core[i] = v[i] / kBohr3ToAng3;
  }
  return u;
}

template <class F>
int guarded_upf(F&& f) {
  try {
    f();
    return PWK_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return PWK_E_OOM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return PWK_E_INVALID;
  }
}

}  // namespace
}  // namespace pwk

extern "C" {

int pwk_upf_parse(const char* text, size_t len, pwk_upf** out) {
  return pwk::guarded_upf([&] {
    PWK_REQUIRE(out != nullptr, PWK_E_INVALID, "null output pointer");
    *out = nullptr;
    PWK_REQUIRE(text != nullptr, PWK_E_INVALID, "null UPF text");
    *out = pwk::parse_body(std::string(text, len)).release();
  });
}

int pwk_upf_load(const char* path, pwk_upf** out) {
  return pwk::guarded_upf([&] {
    PWK_REQUIRE(out != nullptr, PWK_E_INVALID, "null output pointer");
    *out = nullptr;
    PWK_REQUIRE(path != nullptr, PWK_E_INVALID, "null path");
    std::ifstream f(path, std::ios::binary);
    PWK_REQUIRE(f.good(), PWK_E_INVALID, std::string("cannot open ") + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    *out = pwk::parse_body(ss.str()).release();
  });
}

int pwk_upf_species(const pwk_upf* u, pwk_species* s) {
  return pwk::guarded_upf([&] {
    PWK_REQUIRE(u && s, PWK_E_INVALID, "null UPF or species");
    s->atomic_number = u->atomic_number;
    s->z_valence = u->z_valence;
    s->mesh = u->mesh;
    s->r = u->r.data();
    s->rab = u->rab.data();
    s->v_local = u->v_local.data();
    s->nproj = (int)u->proj_l.size();
    s->proj_l = u->proj_l.empty() ? nullptr : u->proj_l.data();
    s->beta = u->beta.empty() ? nullptr : u->beta.data();
    s->dij = u->dij.empty() ? nullptr : u->dij.data();
    s->rho_atom = u->rho_atom.data();
    s->rho_core = u->nlcc ? u->rho_core.data() : nullptr;
  });
}

int pwk_upf_info(const pwk_upf* u, char symbol[4], int* l_max) {
  return pwk::guarded_upf([&] {
    PWK_REQUIRE(u, PWK_E_INVALID, "null UPF");
    if (symbol) {
      std::memset(symbol, 0, 4);
      std::strncpy(symbol, u->element.c_str(), 3);
    }
    if (l_max) *l_max = u->l_max;
  });
}

int pwk_upf_free(pwk_upf* u) {
  delete u;
  return PWK_OK;
}

}  // extern "C"
