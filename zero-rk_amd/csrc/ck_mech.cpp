// CHEMKIN-II reader and reference-layout table builder (host side).
#include "ck_mech.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <set>
#include <sstream>

namespace zrk {
namespace {

constexpr double kRu = 8.314462618e3;              // src/zerork/constants.h:10
constexpr double kCalPerMolTact = 4184.0 / kRu;     // constants.h:19

double atomic_mass(const std::string& el, bool* ok) {
  // IUPAC 2007 masses [kg/kmol] of elements 1-92, the table of
  // src/zerork/atomicMassDB.h; symbols match case-insensitively on their
  // first two characters (element::setElement, src/zerork/element.cpp:61-86)
  static const std::map<std::string, double> tbl = {
      {"H", 1.00794}, {"HE", 4.002602}, {"LI", 6.941}, {"BE", 9.012182}, {"B", 10.811},
      {"C", 12.0107}, {"N", 14.0067}, {"O", 15.9994}, {"F", 18.9984032}, {"NE", 20.1797},
      {"NA", 22.98976928}, {"MG", 24.3050}, {"AL", 26.9815386}, {"SI", 28.0855}, {"P", 30.973762},
      {"S", 32.065}, {"CL", 35.453}, {"AR", 39.948}, {"K", 39.0983}, {"CA", 40.078},
      {"SC", 44.955912}, {"TI", 47.867}, {"V", 50.9415}, {"CR", 51.9961}, {"MN", 54.938045},
      {"FE", 55.845}, {"CO", 58.933195}, {"NI", 58.6934}, {"CU", 63.546}, {"ZN", 65.38},
      {"GA", 69.723}, {"GE", 72.64}, {"AS", 74.92160}, {"SE", 78.96}, {"BR", 79.904},
      {"KR", 83.798}, {"RB", 85.4678}, {"SR", 87.62}, {"Y", 88.90585}, {"ZR", 91.224},
      {"NB", 92.90638}, {"MO", 95.96}, {"TC", 98}, {"RU", 101.07}, {"RH", 102.90550},
      {"PD", 106.42}, {"AG", 107.8682}, {"CD", 112.411}, {"IN", 114.818}, {"SN", 118.710},
      {"SB", 121.760}, {"TE", 127.60}, {"I", 126.90447}, {"XE", 131.293}, {"CS", 132.9054519},
      {"BA", 137.327}, {"LA", 138.90547}, {"CE", 140.116}, {"PR", 140.90765}, {"ND", 144.242},
      {"PM", 145}, {"SM", 150.36}, {"EU", 151.964}, {"GD", 157.25}, {"TB", 158.92535},
      {"DY", 162.500}, {"HO", 164.93032}, {"ER", 167.259}, {"TM", 168.93421}, {"YB", 173.054},
      {"LU", 174.9668}, {"HF", 178.49}, {"TA", 180.94788}, {"W", 183.84}, {"RE", 186.207},
      {"OS", 190.23}, {"IR", 192.217}, {"PT", 195.084}, {"AU", 196.966569}, {"HG", 200.59},
      {"TL", 204.3833}, {"PB", 207.2}, {"BI", 208.98040}, {"PO", 209}, {"AT", 210},
      {"RN", 222}, {"FR", 223}, {"RA", 226}, {"AC", 227}, {"TH", 232.03806},
      {"PA", 231.03588}, {"U", 238.02891},
  };
  std::string key = el.substr(0, 2);
  for (auto& c : key) c = (char)std::toupper((unsigned char)c);
  auto it = tbl.find(key);
  *ok = it != tbl.end();
  return *ok ? it->second : 0.0;
}

std::string upper(std::string s) {
  for (auto& c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

std::string strip_comment(const std::string& s) {
  auto p = s.find('!');
  return p == std::string::npos ? s : s.substr(0, p);
}

std::vector<std::string> tokens(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

bool to_double(std::string s, double* v) {
  for (auto& c : s) {
    if (c == 'D') c = 'E';
    if (c == 'd') c = 'e';
  }
  // trim
  size_t a = s.find_first_not_of(" \t\r");
  if (a == std::string::npos) return false;
  size_t b = s.find_last_not_of(" \t\r");
  s = s.substr(a, b - a + 1);
  char* end = nullptr;
  *v = std::strtod(s.c_str(), &end);
  return end && *end == '\0';
}

struct Section {
  std::string key;
  std::vector<std::string> rest;
  std::vector<std::string> lines;
};

std::vector<Section> split_sections(const std::string& text) {
  std::vector<Section> out;
  std::istringstream is(text);
  std::string raw;
  Section* cur = nullptr;
  while (std::getline(is, raw)) {
    if (!raw.empty() && raw.back() == '\r') raw.pop_back();
    std::string line = strip_comment(raw);
    auto tok = tokens(line);
    if (tok.empty()) continue;
    std::string key = upper(tok[0]);
    if (key == "END") {
      cur = nullptr;
      continue;
    }
    std::string k4 = key.substr(0, 4);
    bool in_thermo_card = cur && cur->key == "THER" && raw.size() >= 80;
    if ((k4 == "ELEM" || k4 == "SPEC" || k4 == "THER" || k4 == "REAC") && !in_thermo_card) {
      out.push_back(Section{k4, std::vector<std::string>(tok.begin() + 1, tok.end()), {}});
      cur = &out.back();
      continue;
    }
    if (cur) cur->lines.push_back(cur->key == "THER" ? raw : line);
  }
  return out;
}

bool parse_thermo(const std::vector<std::string>& lines, std::map<std::string, Species>* db,
                  std::string* err) {
  size_t i = 0;
  double tdef_mid = 1000.0;
  while (i < lines.size() && tokens(strip_comment(lines[i])).empty()) ++i;
  if (i < lines.size()) {
    const std::string& s = lines[i];
    auto tok = tokens(strip_comment(s));
    double a, b, c;
    if (tok.size() >= 3 && to_double(tok[0], &a) && to_double(tok[1], &b) &&
        to_double(tok[2], &c) && !(s.size() >= 80 && s[79] == '1')) {
      tdef_mid = b;
      ++i;
    }
  }
  while (i < lines.size()) {
    const std::string& s = lines[i];
    if (s.size() < 80 || s[79] != '1') {
      ++i;
      continue;
    }
    if (i + 3 >= lines.size()) {
      *err = "truncated thermo card";
      return false;
    }
    Species sp;
    auto nt = tokens(s.substr(0, 16));
    if (nt.empty()) {
      *err = "thermo card without a species name";
      return false;
    }
    sp.name = nt[0];
    auto add_el = [&](std::string sym, const std::string& cnt) {
      sym.erase(std::remove(sym.begin(), sym.end(), ' '), sym.end());
      double n;
      if (sym.empty() || !to_double(cnt, &n)) return;
      int k = (int)std::lround(n);
      if (k != 0) sp.comp.push_back({upper(sym), k});
    };
    for (int k = 0; k < 4; ++k) add_el(s.substr(24 + 5 * k, 2), s.substr(26 + 5 * k, 3));
    double tm;
    sp.tmid = to_double(s.substr(65, 8), &tm) && tm != 0.0 ? tm : tdef_mid;
    if (s.size() >= 78) add_el(s.substr(73, 2), s.substr(75, 3));
    double c[14];
    const std::string* L[3] = {&lines[i + 1], &lines[i + 2], &lines[i + 3]};
    int q = 0;
    for (int li = 0; li < 3; ++li) {
      int ncoef = li < 2 ? 5 : 4;
      for (int k = 0; k < ncoef; ++k) {
        if (L[li]->size() < (size_t)(15 * k + 15) || !to_double(L[li]->substr(15 * k, 15), &c[q])) {
          *err = "bad thermo coefficient for " + sp.name;
          return false;
        }
        ++q;
      }
    }
    for (int k = 0; k < 7; ++k) {
      sp.high[k] = c[k];
      sp.low[k] = c[7 + k];
    }
    if (!db->count(sp.name)) (*db)[sp.name] = sp;
    i += 4;
  }
  return true;
}

std::vector<std::pair<std::string, double>> split_side(const std::string& side) {
  std::vector<std::string> parts;
  std::string cur;
  for (char ch : side) {
    bool nonblank = cur.find_first_not_of(" \t") != std::string::npos;
    if (ch == '+' && nonblank) {
      parts.push_back(cur);
      cur.clear();
    } else {
      cur += ch;
    }
  }
  if (cur.find_first_not_of(" \t") != std::string::npos) parts.push_back(cur);
  std::vector<std::pair<std::string, double>> out;
  for (auto p : parts) {
    size_t a = p.find_first_not_of(" \t"), b = p.find_last_not_of(" \t");
    p = p.substr(a, b - a + 1);
    size_t k = 0;
    while (k < p.size() && (std::isdigit((unsigned char)p[k]) || p[k] == '.')) ++k;
    double coef = 1.0;
    std::string name = p;
    if (k > 0 && k < p.size()) {
      coef = std::atof(p.substr(0, k).c_str());
      name = p.substr(k);
      size_t s2 = name.find_first_not_of(" \t");
      name = name.substr(s2);
    }
    out.push_back({name, coef});
  }
  return out;
}

bool parse_reactions(const std::vector<std::string>& lines, const std::vector<std::string>& units,
                     std::vector<ReactionText>* rx, std::string* err) {
  double conv = kCalPerMolTact;
  for (auto t : units) {
    t = upper(t);
    if (t.rfind("KCAL/MOL", 0) == 0) conv = kCalPerMolTact * 1000.0;
    else if (t.rfind("KJOULES/MOL", 0) == 0) conv = 238.846 * kCalPerMolTact;
    else if (t.rfind("JOULES/MOL", 0) == 0) conv = 238.846 * kCalPerMolTact / 1000.0;
    else if (t.rfind("CAL/MOL", 0) == 0) conv = kCalPerMolTact;
    else if (t.rfind("KELVIN", 0) == 0) conv = 1.0;
    else if (t.rfind("MOLECULES", 0) == 0 || t.rfind("EVOLT", 0) == 0) {
      *err = "unsupported REACTIONS units " + t;
      return false;
    }
  }
  for (const auto& line0 : lines) {
    auto tok = tokens(line0);
    if (tok.empty()) continue;
    bool is_rxn = line0.find('=') != std::string::npos && tok[0].find('/') == std::string::npos;
    if (is_rxn) {
      if (tok.size() < 4) {
        *err = "bad reaction line: " + line0;
        return false;
      }
      ReactionText r;
      double A, n, E;
      if (!to_double(tok[tok.size() - 3], &A) || !to_double(tok[tok.size() - 2], &n) ||
          !to_double(tok[tok.size() - 1], &E)) {
        *err = "bad Arrhenius parameters: " + line0;
        return false;
      }
      r.A = A;
      r.n = n;
      r.E = E * conv;
      std::string body;
      for (size_t k = 0; k + 3 < tok.size(); ++k) body += tok[k] + " ";
      size_t p = body.find("<=>");
      size_t plen = 3;
      if (p == std::string::npos) {
        p = body.find("=>");
        plen = 2;
        if (p != std::string::npos) r.reversible = false;
      }
      if (p == std::string::npos) {
        p = body.find('=');
        plen = 1;
      }
      std::string side[2] = {body.substr(0, p), body.substr(p + plen)};
      for (int si = 0; si < 2; ++si) {
        std::string s = side[si];
        size_t lp = s.find("(+");
        if (lp == std::string::npos) {
          // allow "( +M)"
          for (size_t q = 0; q < s.size(); ++q)
            if (s[q] == '(') {
              size_t q2 = s.find_first_not_of(" ", q + 1);
              if (q2 != std::string::npos && s[q2] == '+') lp = q;
              break;
            }
        }
        if (lp != std::string::npos) {
          size_t rp = s.find(')', lp);
          std::string inner = s.substr(lp + 1, rp - lp - 1);
          inner.erase(std::remove(inner.begin(), inner.end(), ' '), inner.end());
          inner = inner.substr(1);  // drop '+'
          r.falloff = true;
          r.falloff_species = (upper(inner) == "M") ? "" : inner;
          s = s.substr(0, lp) + s.substr(rp + 1);
        }
        auto lst = split_side(s);
        std::vector<std::pair<std::string, double>> keep;
        for (auto& e : lst) {
          if (upper(e.first) == "M") r.third_body = true;
          else keep.push_back(e);
        }
        (si == 0 ? r.reac : r.prod) = keep;
      }
      if (r.falloff) r.third_body = false;
      rx->push_back(r);
      continue;
    }
    if (rx->empty()) {
      *err = "auxiliary data before the first reaction";
      return false;
    }
    ReactionText& r = rx->back();
    // name [/ values /] pairs
    size_t i = 0;
    const std::string& s = line0;
    while (i < s.size()) {
      while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
      if (i >= s.size()) break;
      size_t j = i;
      while (j < s.size() && s[j] != '/' && !std::isspace((unsigned char)s[j])) ++j;
      std::string name = s.substr(i, j - i);
      size_t k = j;
      while (k < s.size() && std::isspace((unsigned char)s[k])) ++k;
      std::vector<double> vals;
      std::vector<std::string> raw;
      bool has_vals = false;
      if (k < s.size() && s[k] == '/') {
        size_t e = s.find('/', k + 1);
        if (e == std::string::npos) {
          *err = "unterminated '/' in: " + s;
          return false;
        }
        has_vals = true;
        raw = tokens(s.substr(k + 1, e - k - 1));
        i = e + 1;
      } else {
        i = j;
      }
      std::string key = upper(name);
      if (key == "FORD" || key == "RORD") {
        // species name and its reaction order (CKParser.cpp:1615-1639)
        double v;
        if (raw.size() != 2 || !to_double(raw[1], &v)) {
          *err = key + " needs a species and an order: " + s;
          return false;
        }
        (key == "FORD" ? r.ford : r.rord)[raw[0]] = v;
        r.real_order = true;
        continue;
      }
      for (auto& t : raw) {
        double v;
        if (!to_double(t, &v)) {
          *err = "bad number in: " + s;
          return false;
        }
        vals.push_back(v);
      }
      if (key == "DUP" || key == "DUPLICATE") {
      } else if (key == "LOW") {
        if (vals.size() != 3) { *err = "LOW needs 3 values"; return false; }
        r.has_low = true;
        r.low[0] = vals[0]; r.low[1] = vals[1]; r.low[2] = vals[2] * conv;
      } else if (key == "TROE") {
        if (vals.size() != 3 && vals.size() != 4) { *err = "TROE needs 3 or 4 values"; return false; }
        if (!r.sri.empty()) { *err = "cannot specify both SRI and TROE"; return false; }
        r.troe = vals;
      } else if (key == "SRI") {
        if (vals.size() < 3 || vals.size() > 5) { *err = "SRI needs 3 to 5 values"; return false; }
        if (!r.troe.empty()) { *err = "cannot specify both SRI and TROE"; return false; }
        r.sri = vals;
      } else if (key == "PLOG") {
        // one line per pressure: p [atm], A, n, E (CKParser.cpp:1643-1669)
        if (vals.size() != 4) { *err = "expected 4 PLOG parameters"; return false; }
        if (vals[0] < 0.0) { *err = "PLOG pressures are required to be >= 0.0"; return false; }
        r.plog.push_back({vals[0], vals[1], vals[2], vals[3] * conv});
      } else if (key == "REV") {
        if (vals.size() != 3) { *err = "REV needs 3 values"; return false; }
        r.has_rev = true;
        r.rev[0] = vals[0]; r.rev[1] = vals[1]; r.rev[2] = vals[2] * conv;
      } else if (key == "HIGH" || key == "LT" || key == "RLT" || key == "CHEB") {
        *err = "unsupported auxiliary keyword " + key;
        return false;
      } else if (has_vals && vals.size() == 1) {
        r.eff[name] = vals[0];
      } else if (!name.empty()) {
        *err = "unrecognised auxiliary data '" + name + "'";
        return false;
      }
    }
  }
  return true;
}

bool read_file(const std::string& path, std::string* text) {
  std::ifstream f(path);
  if (!f) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  *text = ss.str();
  return true;
}

}  // namespace

bool load_mechanism(const std::string& mech_file, const std::string& therm_file,
                    Mechanism* out, std::string* err) {
  Mechanism& M = *out;
  std::string text;
  if (!read_file(mech_file, &text)) {
    *err = "cannot read mechanism file " + mech_file;
    return false;
  }
  std::vector<std::string> rx_lines, units, thermo_lines;
  std::vector<std::string> species_names;
  for (auto& sec : split_sections(text)) {
    if (sec.key == "ELEM") {
      for (auto& t : sec.rest) M.elements.push_back(upper(t.substr(0, t.find('/'))));
      for (auto& l : sec.lines)
        for (auto& t : tokens(l)) M.elements.push_back(upper(t.substr(0, t.find('/'))));
    } else if (sec.key == "SPEC") {
      // a repeated name is dropped, keeping the first declaration
      // (CKParser.cpp:647-657 warns and ignores it)
      auto add = [&](const std::string& t) {
        if (std::find(species_names.begin(), species_names.end(), t) == species_names.end())
          species_names.push_back(t);
      };
      for (auto& t : sec.rest) add(t);
      for (auto& l : sec.lines)
        for (auto& t : tokens(l)) add(t);
    } else if (sec.key == "THER") {
      thermo_lines.insert(thermo_lines.end(), sec.lines.begin(), sec.lines.end());
    } else if (sec.key == "REAC") {
      units = sec.rest;
      rx_lines.insert(rx_lines.end(), sec.lines.begin(), sec.lines.end());
    }
  }
  std::map<std::string, Species> db;
  if (!therm_file.empty()) {
    std::string ttext;
    if (!read_file(therm_file, &ttext)) {
      *err = "cannot read thermo file " + therm_file;
      return false;
    }
    for (auto& sec : split_sections(ttext))
      if (sec.key == "THER" && !parse_thermo(sec.lines, &db, err)) return false;
  }
  if (!thermo_lines.empty() && !parse_thermo(thermo_lines, &db, err)) return false;
  std::map<std::string, Species> dbu;
  for (auto& kv : db) dbu[upper(kv.first)] = kv.second;

  const int nsp = (int)species_names.size();
  if (nsp == 0) {
    *err = "no species";
    return false;
  }
  M.nsp = nsp;
  M.mw.assign(nsp, 0.0);
  M.thermo.assign((size_t)nsp * 16, 0.0);
  std::map<std::string, int> idx;
  for (int j = 0; j < nsp; ++j) {
    const std::string& nm = species_names[j];
    if (idx.count(nm)) {
      *err = "duplicate species " + nm;
      return false;
    }
    idx[nm] = j;
    auto it = db.find(nm);
    const Species* sp = nullptr;
    if (it != db.end()) sp = &it->second;
    else if (dbu.count(upper(nm))) sp = &dbu[upper(nm)];
    if (!sp) {
      *err = "no thermo data for species " + nm;
      return false;
    }
    Species s = *sp;
    s.name = nm;
    double w = 0.0;
    for (auto& e : s.comp) {
      bool ok;
      double am = atomic_mass(e.first, &ok);
      if (!ok) {
        *err = "unknown element " + e.first;
        return false;
      }
      w += e.second * am;
    }
    M.mw[j] = w;
    M.thermo[(size_t)j * 16 + 0] = s.tmid;
    for (int k = 0; k < 7; ++k) {
      M.thermo[(size_t)j * 16 + 1 + k] = s.low[k];
      M.thermo[(size_t)j * 16 + 8 + k] = s.high[k];
    }
    M.species.push_back(s);
  }
  if (!parse_reactions(rx_lines, units, &M.rxns, err)) return false;
  M.nrxn = (int)M.rxns.size();

  // which reactions are non-integer (info_net.cpp:411-470): FORD/RORD, or a
  // stoichiometric coefficient further than 1e-14 from an integer
  constexpr double kStoichTol = 1.0e-14;  // src/zerork/constants.h:32
  std::vector<char> nonint(M.nrxn, 0);
  for (int j = 0; j < M.nrxn; ++j) {
    const auto& r = M.rxns[j];
    bool ni = r.real_order;
    for (const auto* side : {&r.reac, &r.prod})
      for (auto& e : *side) {
        const int k = (int)(e.second + 0.5);
        if (std::fabs(e.second - (double)k) > kStoichTol) ni = true;
      }
    nonint[j] = ni;
    // PLOG exclusions (CKParser.cpp:312-345)
    if (!r.plog.empty()) {
      if (r.falloff || r.has_low || !r.troe.empty() || !r.sri.empty()) {
        *err = "PLOG and falloff reactions (LOW/TROE/SRI) are mutually exclusive";
        return false;
      }
      if (r.third_body) {
        *err = "PLOG and 3rd-body reactions M or (M) are mutually exclusive";
        return false;
      }
      if (r.has_rev && r.rev[0] > 0.0) {
        *err = "PLOG does not support REV parameters being specified";
        return false;
      }
    }
  }

  // step numbering (info_net.cpp:46-99) and expanded participants (:179-300);
  // the steps of non-integer reactions have no expanded lists
  auto expand = [&](const std::vector<std::pair<std::string, double>>& lst,
                    std::vector<int>* outv) -> bool {
    for (auto& e : lst) {
      auto it = idx.find(e.first);
      if (it == idx.end()) {
        *err = "unknown species in reaction: " + e.first;
        return false;
      }
      const int k = (int)(e.second + 0.5);
      for (int q = 0; q < k; ++q) outv->push_back(it->second);
    }
    return true;
  };
  // rate-of-progress exponents of a non-integer reaction, by species name
  // (non_integer_reaction_network.cpp:53-205): the summed stoichiometric
  // coefficients of each side, overwritten or extended by FORD (forward) and
  // RORD (reverse)
  auto rop_map = [&](const std::vector<std::pair<std::string, double>>& side,
                     const std::map<std::string, double>& over, NonIntStep* st) -> bool {
    std::map<std::string, double> mp;
    for (auto& e : side) mp[e.first] += e.second;
    for (auto& e : over) mp[e.first] = e.second;
    for (auto& e : mp) {
      auto it = idx.find(e.first);
      if (it == idx.end()) {
        *err = "unknown species in reaction order: " + e.first;
        return false;
      }
      st->rop_sp.push_back(it->second);
      st->rop_pow.push_back(e.second);
    }
    return true;
  };
  auto nu_list = [&](const std::vector<std::pair<std::string, double>>& side,
                     std::vector<int>* sp, std::vector<double>* nu) -> bool {
    for (auto& e : side) {
      auto it = idx.find(e.first);
      if (it == idx.end()) {
        *err = "unknown species in reaction: " + e.first;
        return false;
      }
      sp->push_back(it->second);
      nu->push_back(e.second);
    }
    return true;
  };
  M.reac_ptr.assign(1, 0);
  M.prod_ptr.assign(1, 0);
  M.fwd_step.assign(M.nrxn, -1);
  M.rev_step.assign(M.nrxn, -1);
  std::vector<double> step_order;  // real order (sum of rate-of-progress exponents)
  for (int j = 0; j < M.nrxn; ++j) {
    const auto& r = M.rxns[j];
    std::vector<int> re, pr;
    if (!nonint[j] && (!expand(r.reac, &re) || !expand(r.prod, &pr))) return false;
    const bool has_rev = r.reversible && !(r.has_rev && r.rev[0] == 0.0);
    for (int dir = 0; dir < (has_rev ? 2 : 1); ++dir) {
      const int s = (int)M.step_rxn.size();
      (dir == 0 ? M.fwd_step : M.rev_step)[j] = s;
      M.step_rxn.push_back(j);
      M.step_dir.push_back(dir == 0 ? 1 : -1);
      const std::vector<int>& a = dir == 0 ? re : pr;
      const std::vector<int>& b = dir == 0 ? pr : re;
      M.reac_sp.insert(M.reac_sp.end(), a.begin(), a.end());
      M.reac_ptr.push_back((int)M.reac_sp.size());
      M.prod_sp.insert(M.prod_sp.end(), b.begin(), b.end());
      M.prod_ptr.push_back((int)M.prod_sp.size());
      if (nonint[j]) {
        NonIntStep st;
        st.step = s;
        NonIntStep other;
        if (!rop_map(r.reac, r.ford, dir == 0 ? &st : &other) ||
            !rop_map(r.prod, r.rord, dir == 0 ? &other : &st) ||
            !nu_list(dir == 0 ? r.reac : r.prod, &st.re_sp, &st.re_nu) ||
            !nu_list(dir == 0 ? r.prod : r.reac, &st.pr_sp, &st.pr_nu))
          return false;
        double o = 0.0;
        for (double q : st.rop_pow) o += q;
        step_order.push_back(o);
        M.ni.push_back(std::move(st));
      } else {
        step_order.push_back((double)a.size());
      }
    }
  }
  M.nstep = (int)M.step_rxn.size();
  M.step_ni.assign(M.nstep, -1);
  for (size_t q = 0; q < M.ni.size(); ++q) M.step_ni[M.ni[q].step] = (int)q;

  // Arrhenius parameters with unit conversion (rate_const.cpp:212-260) and
  // merging of triples equal within 1e-6 relative (:275-332, 1093-1128);
  // PLOG forward steps have no Arrhenius entry (:159-190)
  struct Arr { double A, n, Ta; int step; };
  std::vector<Arr> arr;
  M.rev_type.assign(M.nrxn, kRevNone);
  M.keq_ndelta.assign(M.nrxn, 0.0);
  M.rxn_plog.assign(M.nrxn, -1);
  for (int j = 0; j < M.nrxn; ++j) {
    const auto& r = M.rxns[j];
    const int s = M.fwd_step[j];
    if (r.plog.empty()) {
      double A = r.A * std::pow(1000.0, 1.0 - step_order[s]);
      if (r.third_body) A /= 1000.0;
      arr.push_back({A, r.n, r.E, s});
    } else {
      // PLOG lines (rate_const.cpp:1146-1186; plog_reaction.cpp:10-135):
      // selection sort by pressure, then lines within 1e-8 relative of the
      // previous distinct pressure share it
      std::vector<std::array<double, 4>> ln = r.plog;
      const double ac = std::pow(1000.0, 1.0 - step_order[s]);
      for (auto& l : ln) {
        l[0] *= 1.01325e5;
        l[1] *= ac;
      }
      for (size_t a = 0; a + 1 < ln.size(); ++a) {
        size_t mi = a;
        for (size_t b = a + 1; b < ln.size(); ++b)
          if (ln[b][0] < ln[mi][0]) mi = b;
        if (mi != a) std::swap(ln[a], ln[mi]);
      }
      PlogStep ps;
      ps.step = s;
      for (size_t a = 0; a < ln.size(); ++a) {
        if (a == 0 || std::fabs(ln[a][0] - ps.p.back()) > std::fabs(ps.p.back() * 1.0e-8)) {
          ps.p.push_back(ln[a][0]);
          ps.line_ptr.push_back((int)a);
        }
        ps.lnA.push_back(std::log(std::fabs(ln[a][1])));
        ps.sign.push_back(ln[a][1] < 0.0 ? -1.0 : 1.0);
        ps.beta.push_back(ln[a][2]);
        ps.Ta.push_back(ln[a][3]);
      }
      ps.line_ptr.push_back((int)ln.size());
      for (double pp : ps.p) ps.lnp.push_back(std::log(pp));
      M.rxn_plog[j] = (int)M.plog.size();
      M.plog.push_back(std::move(ps));
    }
    int sr = M.rev_step[j];
    if (sr >= 0) {
      if (r.has_rev && r.rev[0] > 0.0) {
        double Ar = r.rev[0] * std::pow(1000.0, 1.0 - step_order[sr]);
        if (r.third_body) Ar /= 1000.0;
        arr.push_back({Ar, r.rev[1], r.rev[2], sr});
        M.rev_type[j] = kRevArrhenius;
      } else {
        M.rev_type[j] = kRevKeq;
        M.keq_ndelta[j] = nonint[j] ? step_order[sr] - step_order[s]
                                    : (double)((M.prod_ptr[s + 1] - M.prod_ptr[s]) -
                                               (M.reac_ptr[s + 1] - M.reac_ptr[s]));
      }
    }
  }
  std::vector<Arr> sorted = arr;
  std::sort(sorted.begin(), sorted.end(), [](const Arr& x, const Arr& y) {
    if (x.A != y.A) return x.A < y.A;
    if (x.Ta != y.Ta) return x.Ta > y.Ta;
    return x.n < y.n;
  });
  auto same = [](const Arr& x, const Arr& y) {
    const double a[3][2] = {{x.A, y.A}, {x.n, y.n}, {x.Ta, y.Ta}};
    for (auto& p : a) {
      double d = std::fabs(p[0] - p[1]);
      if (d > 1.0e-6 * std::fabs(0.5 * (p[0] + p[1])) && d > 0.0) return false;
    }
    return true;
  };
  M.step_has_arr.assign(M.nstep, 0);
  M.step_lnA.assign(M.nstep, 0.0);
  M.step_beta.assign(M.nstep, 0.0);
  M.step_Ta.assign(M.nstep, 0.0);
  const Arr* last = nullptr;
  for (const auto& e : sorted) {
    if (!last || !same(*last, e)) last = &e;
    M.step_has_arr[e.step] = 1;
    M.step_lnA[e.step] = std::log(last->A);
    M.step_beta[e.step] = last->n;
    M.step_Ta[e.step] = last->Ta;
  }

  // third-body / falloff enhanced lists, efficiency - 1, species-name order
  M.tb_sp_ptr.assign(1, 0);
  M.falloff_type.assign(M.nrxn, -1);
  M.falloff_spc.assign(M.nrxn, -1);
  M.falloff_param.assign((size_t)M.nrxn * kFoParams, 0.0);
  for (int j = 0; j < M.nrxn; ++j) {
    const auto& r = M.rxns[j];
    bool single = r.falloff && !r.falloff_species.empty();
    if ((r.third_body || r.falloff) && !single) {
      for (auto& kv : r.eff) {
        auto it = idx.find(kv.first);
        if (it == idx.end()) continue;  // reference ignores unknown species
        M.tb_sp.push_back(it->second);
        M.tb_eff.push_back(kv.second - 1.0);
      }
    }
    M.tb_sp_ptr.push_back((int)M.tb_sp.size());
    if (r.falloff) {
      if (!r.has_low) {
        *err = "falloff reaction without LOW";
        return false;
      }
      if (single) {
        auto it = idx.find(r.falloff_species);
        if (it == idx.end()) {
          *err = "unknown falloff species " + r.falloff_species;
          return false;
        }
        M.falloff_spc[j] = it->second;
      }
      const int s = M.fwd_step[j];
      double* p = &M.falloff_param[(size_t)j * kFoParams];
      p[0] = std::log(r.low[0] * std::pow(1000.0, -step_order[s]));
      p[1] = r.low[1];
      p[2] = r.low[2];
      if (!r.sri.empty()) {
        // a, b, 1/c (or -1 when c <= 0), d (default 1), e (default 0)
        M.falloff_type[j] = kSri;
        p[3] = r.sri[0];
        p[4] = r.sri[1];
        p[5] = r.sri[2] > 0.0 ? 1.0 / r.sri[2] : -1.0;
        p[6] = r.sri.size() >= 4 ? r.sri[3] : 1.0;
        p[7] = r.sri.size() == 5 ? r.sri[4] : 0.0;
      } else if (r.troe.empty()) {
        M.falloff_type[j] = kLindemann;
      } else {
        for (size_t k = 0; k < r.troe.size(); ++k) p[3 + k] = r.troe[k];
        M.falloff_type[j] = r.troe.size() == 3 ? kTroe3 : kTroe4;
      }
    }
  }
  return true;
}

}  // namespace zrk
