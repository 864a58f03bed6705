#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <unordered_map>

#include "prep.hpp"

namespace nir {

// ------------------------------------------------------------------ checker
CheckReport verify_solution(const Model& model, const std::vector<double>& x, double claimed_obj,
                            double tol) {
    CheckReport rep;
    if (static_cast<Idx>(x.size()) != model.ncol) {
        rep.passed = false;
        rep.worst_item = "Solution size mismatch";
        return rep;
    }
    // Track the largest violation; name the item only when it exceeds tol.
    auto note = [&](double viol, double& worst, const char* what, const std::string& nm) {
        if (viol > worst) {
            worst = viol;
            if (viol > tol) rep.worst_item = what + nm;
        }
    };
    for (Idx j = 0; j < model.ncol; ++j) {
        const double v = x[j];
        const std::string& nm = model.col_names[j];
        if (!std::isfinite(v)) {
            rep.passed = false;
            rep.worst_item = "NaN/Inf in variable " + nm;
            return rep;
        }
        if (finite_bound(model.collo[j]))
            note(amax(0.0, model.collo[j] - v) / (1.0 + std::fabs(model.collo[j])), rep.max_bound_viol,
                 "Lower bound on ", nm);
        if (finite_bound(model.colup[j]))
            note(amax(0.0, v - model.colup[j]) / (1.0 + std::fabs(model.colup[j])), rep.max_bound_viol,
                 "Upper bound on ", nm);
        if (model.vartype[j] == VarType::Integer || model.vartype[j] == VarType::Binary)
            note(std::fabs(v - std::round(v)), rep.max_int_viol, "Integrality on ", nm);
    }
    // Row activities in double-double, accumulated column by column.
    std::vector<DD> act(static_cast<size_t>(model.nrow));
    for (Idx j = 0; j < model.ncol; ++j) {
        const double xj = x[j];
        if (xj == 0.0) continue;
        for (Off p = model.A.start[j]; p < model.A.start[j + 1]; ++p)
            act[model.A.index[p]].add_prod(model.A.value[p], xj);
    }
    for (Idx i = 0; i < model.nrow; ++i) {
        const double r = act[i].value();
        const std::string& nm = model.row_names[i];
        if (finite_bound(model.rowlo[i]))
            note(amax(0.0, model.rowlo[i] - r) / (1.0 + std::fabs(model.rowlo[i])), rep.max_row_viol,
                 "Row lower bound on ", nm);
        if (finite_bound(model.rowup[i]))
            note(amax(0.0, r - model.rowup[i]) / (1.0 + std::fabs(model.rowup[i])), rep.max_row_viol,
                 "Row upper bound on ", nm);
    }
    rep.computed_obj = model.eval_objective(x);
    if (std::isfinite(claimed_obj))
        rep.obj_rel_err = std::fabs(rep.computed_obj - claimed_obj) / (1.0 + std::fabs(rep.computed_obj));
    if (rep.max_bound_viol > tol || rep.max_row_viol > tol || rep.max_int_viol > tol || rep.obj_rel_err > 1e-5)
        rep.passed = false;
    return rep;
}

// ------------------------------------------------------------------ .sol files
bool write_sol(const std::string& path, const Model& m, Status st, double obj, double gap,
               const std::vector<double>& x) {
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "# status %s\n", status_name(st));
    std::fprintf(f, "# objective %.17g\n", obj);
    std::fprintf(f, "# gap %.6e\n", gap);
    for (Idx j = 0; j < m.ncol && j < static_cast<Idx>(x.size()); ++j) {
        const std::string nm = j < static_cast<Idx>(m.col_names.size()) ? m.col_names[j] : "C" + std::to_string(j);
        std::fprintf(f, "%s %.17g\n", nm.c_str(), x[j]);
    }
    std::fclose(f);
    return true;
}

bool read_sol(const std::string& path, const Model& m, Status& st, double& obj, std::vector<double>& x) {
    std::ifstream in(path);
    if (!in.is_open()) return false;
    x.assign(static_cast<size_t>(m.ncol), 0.0);
    st = Status::Optimal;
    obj = 0.0;
    std::unordered_map<std::string, Idx> by_name;
    for (Idx j = 0; j < m.ncol; ++j) by_name[m.col_names[j]] = j;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::istringstream ss(line);
        if (line[0] == '#') {
            std::string hash, key, val;
            ss >> hash >> key >> val;
            if (key == "objective") {
                try { obj = std::stod(val); } catch (const std::exception&) { return false; }
            }
            continue;
        }
        std::string nm;
        double v = 0.0;
        if (ss >> nm >> v) {
            auto it = by_name.find(nm);
            if (it != by_name.end()) x[it->second] = v;
        }
    }
    return true;
}

// ------------------------------------------------------------------ MPS
namespace {

std::string upper(std::string s) {
    for (char& c : s) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
    return s;
}

// Fortran-style exponents ("1.5D3") are accepted.
double mps_number(std::string s) {
    for (char& c : s)
        if (c == 'D' || c == 'd') c = 'E';
    return std::stod(s);
}

enum class Part { None, Name, ObjSense, ObjName, Rows, Columns, Rhs, Ranges, Bounds, QuadObj, QMatrix, Sos, End };
enum class RowKind { L, G, E };

}  // namespace

Status read_mps(const std::string& path, Model& out, std::string& err) {
    std::ifstream in(path);
    if (!in.is_open()) {
        err = "Cannot open MPS file: " + path;
        return Status::InputError;
    }
    try {
        Model m;
        m.name = "mps_model";
        std::string obj_row;
        bool sense_pending = false;
        Part part = Part::None;
        std::unordered_map<std::string, Idx> row_id, col_id;
        std::vector<RowKind> kind;
        std::vector<double> rhs, range;
        std::vector<char> has_range;
        std::vector<Entry> entries, qentries;
        bool int_block = false;

        auto column = [&](const std::string& nm) -> Idx {
            auto it = col_id.find(nm);
            if (it != col_id.end()) return it->second;
            const Idx j = m.ncol++;
            col_id[nm] = j;
            m.col_names.push_back(nm);
            m.cost.push_back(0.0);
            m.collo.push_back(0.0);
            m.colup.push_back(kInf);
            m.vartype.push_back(int_block ? VarType::Integer : VarType::Continuous);
            return j;
        };

        std::string line;
        int lineno = 0;
        while (std::getline(in, line)) {
            ++lineno;
            if (!line.empty() && line.back() == '\r') line.pop_back();
            if (line.empty() || line[0] == '*') continue;
            std::istringstream ss(line);
            std::vector<std::string> tok;
            for (std::string t; ss >> t;) tok.push_back(t);
            if (tok.empty()) continue;
            const std::string head = upper(tok[0]);

            // Section headers start in column 0.
            if (!std::isspace(static_cast<unsigned char>(line[0]))) {
                bool header = true;
                if (head == "NAME") {
                    if (tok.size() >= 2) m.name = tok[1];
                    part = Part::Name;
                } else if (head == "OBJSENSE") {
                    if (tok.size() >= 2) {
                        m.sense = upper(tok[1]).rfind("MAX", 0) == 0 ? Sense::Maximize : Sense::Minimize;
                    } else {
                        sense_pending = true;
                        part = Part::ObjSense;
                    }
                } else if (head == "OBJNAME") {
                    if (tok.size() >= 2) obj_row = tok[1];
                    else part = Part::ObjName;
                } else if (head == "ROWS") part = Part::Rows;
                else if (head == "COLUMNS") part = Part::Columns;
                else if (head == "RHS") part = Part::Rhs;
                else if (head == "RANGES") part = Part::Ranges;
                else if (head == "BOUNDS") part = Part::Bounds;
                else if (head == "QUADOBJ" || head == "QSECTION") part = Part::QuadObj;
                else if (head == "QMATRIX") part = Part::QMatrix;
                else if (head == "SOS") part = Part::Sos;
                else if (head == "ENDATA") { part = Part::End; break; }
                else header = false;
                if (header) continue;
            }

            if (part == Part::ObjSense && sense_pending) {
                m.sense = head.rfind("MAX", 0) == 0 ? Sense::Maximize : Sense::Minimize;
                sense_pending = false;
                continue;
            }
            if (part == Part::ObjName && obj_row.empty()) {
                obj_row = tok[0];
                continue;
            }
            switch (part) {
                case Part::Rows: {
                    if (tok.size() < 2) break;
                    const char t = static_cast<char>(std::toupper(static_cast<unsigned char>(tok[0][0])));
                    const std::string& rn = tok[1];
                    if (t == 'N') {
                        if (obj_row.empty()) obj_row = rn;   // further free rows are ignored
                        break;
                    }
                    if (row_id.count(rn)) {
                        err = "Duplicate row name at line " + std::to_string(lineno);
                        return Status::InputError;
                    }
                    row_id[rn] = m.nrow++;
                    m.row_names.push_back(rn);
                    kind.push_back(t == 'L' ? RowKind::L : t == 'G' ? RowKind::G : RowKind::E);
                    rhs.push_back(0.0);
                    range.push_back(0.0);
                    has_range.push_back(0);
                    break;
                }
                case Part::Columns: {
                    bool marker = false;
                    for (const std::string& t : tok) {
                        if (t.find("INTORG") != std::string::npos) { int_block = true; marker = true; }
                        else if (t.find("INTEND") != std::string::npos) { int_block = false; marker = true; }
                    }
                    if (marker || tok.size() < 3) break;
                    const Idx j = column(tok[0]);
                    for (size_t k = 1; k + 1 < tok.size(); k += 2) {
                        const double v = mps_number(tok[k + 1]);
                        if (tok[k] == obj_row) {
                            m.cost[j] += v;
                        } else {
                            auto it = row_id.find(tok[k]);
                            if (it != row_id.end()) entries.push_back({it->second, j, v});
                        }
                    }
                    break;
                }
                case Part::Rhs:
                case Part::Ranges: {
                    // An odd token count means the first token names the vector.
                    for (size_t k = tok.size() % 2; k + 1 < tok.size(); k += 2) {
                        const double v = mps_number(tok[k + 1]);
                        if (part == Part::Rhs && tok[k] == obj_row) {
                            m.obj_offset = -v;   // objective RHS is minus the constant
                            continue;
                        }
                        auto it = row_id.find(tok[k]);
                        if (it == row_id.end()) continue;
                        if (part == Part::Rhs) rhs[it->second] = v;
                        else { range[it->second] = v; has_range[it->second] = 1; }
                    }
                    break;
                }
                case Part::Bounds: {
                    if (tok.size() < 3) break;
                    const std::string bt = upper(tok[0]);
                    const bool valued = bt == "LO" || bt == "UP" || bt == "FX" || bt == "LI" || bt == "UI" || bt == "SC";
                    std::string cn;
                    double v = 0.0;
                    if (valued) {
                        cn = tok.size() >= 4 ? tok[2] : tok[1];
                        v = mps_number(tok.size() >= 4 ? tok[3] : tok[2]);
                    } else {
                        cn = tok.size() >= 3 ? tok[2] : tok[1];
                    }
                    auto it = col_id.find(cn);
                    if (it == col_id.end()) {
                        err = "Unknown column in BOUNDS: " + cn;
                        return Status::InputError;
                    }
                    const Idx j = it->second;
                    if (bt == "LO") m.collo[j] = v;
                    else if (bt == "UP") {
                        if (v < 0.0 && m.collo[j] == 0.0) m.collo[j] = -kInf;   // legacy negative-UP rule
                        m.colup[j] = v;
                    } else if (bt == "FX") { m.collo[j] = v; m.colup[j] = v; }
                    else if (bt == "FR") { m.collo[j] = -kInf; m.colup[j] = kInf; }
                    else if (bt == "MI") m.collo[j] = -kInf;
                    else if (bt == "PL") m.colup[j] = kInf;
                    else if (bt == "BV") { m.vartype[j] = VarType::Binary; m.collo[j] = 0.0; m.colup[j] = 1.0; }
                    else if (bt == "LI") { m.vartype[j] = VarType::Integer; m.collo[j] = v; }
                    else if (bt == "UI") { m.vartype[j] = VarType::Integer; m.colup[j] = v; }
                    break;
                }
                case Part::QuadObj:
                case Part::QMatrix: {
                    if (tok.size() < 3) break;
                    auto a = col_id.find(tok[0]), b = col_id.find(tok[1]);
                    if (a == col_id.end() || b == col_id.end()) break;
                    Idx i = a->second, j = b->second;
                    const double v = mps_number(tok[2]);
                    if (part == Part::QMatrix) {
                        if (i <= j) qentries.push_back({i, j, v});
                    } else {
                        if (i > j) std::swap(i, j);
                        qentries.push_back({i, j, v});
                    }
                    break;
                }
                default: break;
            }
        }

        m.rowlo.resize(static_cast<size_t>(m.nrow));
        m.rowup.resize(static_cast<size_t>(m.nrow));
        for (Idx i = 0; i < m.nrow; ++i) {
            const double b = rhs[i];
            if (!has_range[i]) {
                if (kind[i] == RowKind::L) { m.rowlo[i] = -kInf; m.rowup[i] = b; }
                else if (kind[i] == RowKind::G) { m.rowlo[i] = b; m.rowup[i] = kInf; }
                else { m.rowlo[i] = b; m.rowup[i] = b; }
                continue;
            }
            const double r = range[i];
            if (kind[i] == RowKind::G) { m.rowlo[i] = b; m.rowup[i] = b + std::fabs(r); }
            else if (kind[i] == RowKind::L) { m.rowlo[i] = b - std::fabs(r); m.rowup[i] = b; }
            else if (r >= 0.0) { m.rowlo[i] = b; m.rowup[i] = b + r; }
            else { m.rowlo[i] = b + r; m.rowup[i] = b; }
        }
        m.A = CscMatrix::assemble(m.nrow, m.ncol, entries);
        m.Q = CscMatrix::assemble(m.ncol, m.ncol, qentries);
        out = std::move(m);
        return out.validate(err);
    } catch (const std::exception& e) {
        err = std::string("Malformed number in MPS file: ") + e.what();
        return Status::InputError;
    }
}

bool write_mps(const std::string& path, const Model& m) {
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    auto rname = [&](Idx i) { return i < static_cast<Idx>(m.row_names.size()) ? m.row_names[i] : "R" + std::to_string(i); };
    auto cname = [&](Idx j) { return j < static_cast<Idx>(m.col_names.size()) ? m.col_names[j] : "C" + std::to_string(j); };
    std::fprintf(f, "NAME          %s\n", m.name.c_str());
    std::fprintf(f, "OBJSENSE\n  %s\n", m.sense == Sense::Maximize ? "MAX" : "MIN");
    std::fprintf(f, "ROWS\n N  OBJ\n");
    for (Idx i = 0; i < m.nrow; ++i) {
        const bool lf = finite_bound(m.rowlo[i]), uf = finite_bound(m.rowup[i]);
        char t = 'E';
        if (!lf && uf) t = 'L';
        else if (lf && !uf) t = 'G';
        else if (m.rowlo[i] < m.rowup[i]) t = 'G';   // ranged: G row plus RANGES entry
        std::fprintf(f, " %c  %s\n", t, rname(i).c_str());
    }
    std::fprintf(f, "COLUMNS\n");
    bool in_int = false;
    for (Idx j = 0; j < m.ncol; ++j) {
        const bool is_int = m.vartype[j] == VarType::Integer || m.vartype[j] == VarType::Binary;
        if (is_int != in_int) {
            std::fprintf(f, is_int ? "    MARK0000  'MARKER'                 'INTORG'\n"
                                   : "    MARK0001  'MARKER'                 'INTEND'\n");
            in_int = is_int;
        }
        const std::string cn = cname(j);
        if (m.cost[j] != 0.0) std::fprintf(f, "    %-10s %-10s %.17g\n", cn.c_str(), "OBJ", m.cost[j]);
        for (Off p = m.A.start[j]; p < m.A.start[j + 1]; ++p)
            std::fprintf(f, "    %-10s %-10s %.17g\n", cn.c_str(), rname(m.A.index[p]).c_str(), m.A.value[p]);
    }
    if (in_int) std::fprintf(f, "    MARK0001  'MARKER'                 'INTEND'\n");
    std::fprintf(f, "RHS\n");
    if (m.obj_offset != 0.0) std::fprintf(f, "    RHS1       %-10s %.17g\n", "OBJ", -m.obj_offset);
    for (Idx i = 0; i < m.nrow; ++i) {
        const double b = finite_bound(m.rowlo[i]) ? m.rowlo[i] : m.rowup[i];
        if (b != 0.0 && finite_bound(b)) std::fprintf(f, "    RHS1       %-10s %.17g\n", rname(i).c_str(), b);
    }
    std::fprintf(f, "RANGES\n");
    for (Idx i = 0; i < m.nrow; ++i)
        if (finite_bound(m.rowlo[i]) && finite_bound(m.rowup[i]) && m.rowlo[i] < m.rowup[i])
            std::fprintf(f, "    RNG1       %-10s %.17g\n", rname(i).c_str(), m.rowup[i] - m.rowlo[i]);
    std::fprintf(f, "BOUNDS\n");
    for (Idx j = 0; j < m.ncol; ++j) {
        const std::string cn = cname(j);
        const bool lf = finite_bound(m.collo[j]), uf = finite_bound(m.colup[j]);
        if (m.vartype[j] == VarType::Binary && m.collo[j] == 0.0 && m.colup[j] == 1.0) {
            std::fprintf(f, " BV BND1       %s\n", cn.c_str());
            continue;
        }
        if (!lf && !uf) { std::fprintf(f, " FR BND1       %s\n", cn.c_str()); continue; }
        if (lf && uf && m.collo[j] == m.colup[j]) {
            std::fprintf(f, " FX BND1       %-10s %.17g\n", cn.c_str(), m.collo[j]);
            continue;
        }
        if (!lf) std::fprintf(f, " MI BND1       %s\n", cn.c_str());
        else if (m.collo[j] != 0.0) std::fprintf(f, " LO BND1       %-10s %.17g\n", cn.c_str(), m.collo[j]);
        if (uf) std::fprintf(f, " UP BND1       %-10s %.17g\n", cn.c_str(), m.colup[j]);
    }
    if (m.Q.nnz() > 0) {
        std::fprintf(f, "QUADOBJ\n");
        for (Idx j = 0; j < m.Q.ncol; ++j)
            for (Off p = m.Q.start[j]; p < m.Q.start[j + 1]; ++p)
                std::fprintf(f, "    %-10s %-10s %.17g\n", cname(m.Q.index[p]).c_str(), cname(j).c_str(), m.Q.value[p]);
    }
    std::fprintf(f, "ENDATA\n");
    std::fclose(f);
    return true;
}

}  // namespace nir
