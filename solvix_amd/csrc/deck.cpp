// Reader for the Abaqus-style .inp subset of the linear-static path, for the
// request server's host workers (svx_server_submit_deck) and svx_deck_parse.
// Semantics follow crates/solvix-io/src/parser.rs for the keywords the path's
// decks use, as solvix_amd/inp.py restates them:
//   *NODE            id, x, y[, z]; deck order; z = 0 for 2-field records (146-173)
//   *ELEMENT         TYPE=CPS3|CPS4|C3D8 [, ELSET=]; records continue over
//                    lines ending in a comma; other types are skipped (175-256)
//   *NSET / *ELSET   [, GENERATE]; names upper-cased; members appended, duplicates
//                    kept; an element set named twice is one set (258-312)
//   *MATERIAL NAME=  then *ELASTIC E, nu (first data line) (496-535)
//   *SOLID SECTION   ELSET=, MATERIAL=; the thickness is model-global (537-584)
//   *STEP [OP=NEW]   starts from the previous step's state, or empty (1114-1150)
//   *BOUNDARY        node|set, dof[, dof2] (742-796)
//   *CLOAD           node|set, dof, value, expanded over the set (798-844)
// The model's loads and constraints are the last step's, or those given
// outside any step (1325-1336). Other keywords and their data are skipped.
#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <unordered_map>

#include "api_internal.hpp"
#include "deck.hpp"

namespace svx {
namespace {

struct ParseFail {
    int64_t line;
    std::string msg;
};

std::string trim(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}

std::string upper(std::string s) {
    for (char& c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}

std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

std::vector<std::string> split_commas(const std::string& s) {
    std::vector<std::string> out;
    size_t b = 0;
    for (;;) {
        size_t c = s.find(',', b);
        out.push_back(trim(s.substr(b, c == std::string::npos ? std::string::npos : c - b)));
        if (c == std::string::npos) break;
        b = c + 1;
    }
    return out;
}

// Comma-separated fields, trailing empty fields dropped.
std::vector<std::string> fields(const std::string& s) {
    std::vector<std::string> f = split_commas(s);
    while (!f.empty() && f.back().empty()) f.pop_back();
    return f;
}

int64_t to_int(const std::string& s, int64_t line) {
    errno = 0;
    char* end = nullptr;
    long long v = std::strtoll(s.c_str(), &end, 10);
    if (s.empty() || errno || !end || *end != '\0') throw ParseFail{line, "invalid integer '" + s + "'"};
    return v;
}

double to_real(const std::string& s, int64_t line) {
    char* end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (s.empty() || !end || *end != '\0') throw ParseFail{line, "invalid number '" + s + "'"};
    return v;
}

uint32_t to_id(const std::string& s, int64_t line) {
    int64_t v = to_int(s, line);
    if (v < 0 || v > (int64_t)UINT32_MAX) throw ParseFail{line, "id out of range '" + s + "'"};
    return (uint32_t)v;
}

struct Line {
    int64_t no;
    bool keyword;
    std::string text;
};

struct LoadSet {
    std::vector<std::pair<std::vector<uint32_t>, std::vector<int32_t>>> bcs;
    std::vector<std::tuple<uint32_t, int32_t, double>> cloads;
};

void read(const char* text, size_t len, OwnedModel& M) {
    // lines: \n, \r\n and \r end a line; blank lines and ** comments go
    std::vector<Line> items;
    {
        int64_t no = 0;
        size_t i = 0;
        while (i < len) {
            size_t j = i;
            while (j < len && text[j] != '\n' && text[j] != '\r') ++j;
            ++no;
            std::string s = trim(std::string(text + i, j - i));
            if (!s.empty() && s.compare(0, 2, "**") != 0) items.push_back({no, s[0] == '*', s});
            if (j < len && text[j] == '\r' && j + 1 < len && text[j + 1] == '\n') ++j;
            i = j + 1;
        }
    }
    std::unordered_map<std::string, std::vector<uint32_t>> nsets;
    std::vector<std::pair<std::string, std::vector<uint32_t>>> elsets;
    std::vector<std::string> mat_names;
    std::vector<std::pair<std::string, std::string>> assignments;
    std::vector<std::vector<uint32_t>> conn;
    std::string pending_material;
    bool have_pending = false;
    LoadSet model_state;
    std::vector<LoadSet> steps;
    std::unique_ptr<LoadSet> cur;

    auto upsert_elset = [&](const std::string& nm, const std::vector<uint32_t>& ids) {
        const std::string key = upper(nm);
        for (auto& s : elsets)
            if (s.first == key) {
                s.second.insert(s.second.end(), ids.begin(), ids.end());
                return;
            }
        elsets.emplace_back(key, ids);
    };
    auto resolve = [&](const std::string& target, int64_t line) -> std::vector<uint32_t> {
        auto it = nsets.find(upper(target));
        if (it != nsets.end()) return it->second;
        try {
            return {to_id(target, line)};
        } catch (const ParseFail&) {
            throw ParseFail{line, "unknown node set or node id '" + target + "'"};
        }
    };

    size_t j = 0;
    while (j < items.size()) {
        if (!items[j].keyword) {
            ++j;
            continue;
        }
        // keyword line: name, then KEY=value / FLAG parameters
        const std::vector<std::string> parts = split_commas(items[j].text.substr(1));
        const std::string kw = lower(parts[0]);
        std::unordered_map<std::string, std::string> params;
        for (size_t k = 1; k < parts.size(); ++k) {
            if (parts[k].empty()) continue;
            size_t eq = parts[k].find('=');
            if (eq == std::string::npos)
                params[lower(parts[k])] = "";
            else
                params[lower(trim(parts[k].substr(0, eq)))] = trim(parts[k].substr(eq + 1));
        }
        auto param = [&](const char* key, std::string* v) {
            auto it = params.find(key);
            if (it == params.end()) return false;
            *v = it->second;
            return true;
        };
        ++j;
        std::vector<const Line*> data;
        while (j < items.size() && !items[j].keyword) data.push_back(&items[j++]);

        if (kw == "node") {
            for (const Line* d : data) {
                std::vector<std::string> f = fields(d->text);
                if (f.size() != 3 && f.size() != 4)
                    throw ParseFail{d->no, "*NODE expects id, x, y[, z]; got " + std::to_string(f.size()) + " field(s)"};
                M.node_ids.push_back(to_id(f[0], d->no));
                M.coords.push_back(to_real(f[1], d->no));
                M.coords.push_back(to_real(f[2], d->no));
                M.coords.push_back(f.size() == 4 ? to_real(f[3], d->no) : 0.0);
            }
        } else if (kw == "element") {
            std::string ty;
            param("type", &ty);
            ty = upper(ty);
            int kind = ty == "CPS3" ? SVX_CPS3 : ty == "CPS4" ? SVX_CPS4 : ty == "C3D8" ? SVX_C3D8 : -1;
            if (kind < 0) continue;
            const size_t want = 1 + (size_t)nodes_of((uint8_t)kind);
            std::string es;
            const bool has_es = param("elset", &es);
            size_t q = 0;
            while (q < data.size()) {
                std::vector<std::string> f;
                const int64_t first = data[q]->no;
                while (q < data.size()) {
                    const std::string& t = data[q]->text;
                    const bool ends_comma = !t.empty() && t.back() == ',';
                    for (std::string& s : split_commas(t)) f.push_back(s);
                    ++q;
                    size_t filled = 0;
                    for (const std::string& s : f) filled += !s.empty();
                    if (!ends_comma || filled >= want) break;
                }
                std::vector<std::string> g;
                for (std::string& s : f)
                    if (!s.empty()) g.push_back(s);
                if (g.size() != want)
                    throw ParseFail{first, "*ELEMENT expects " + std::to_string(want) + " field(s); got " +
                                               std::to_string(g.size())};
                const uint32_t eid = to_id(g[0], first);
                M.elem_ids.push_back(eid);
                M.kinds.push_back((uint8_t)kind);
                std::vector<uint32_t> c;
                for (size_t k = 1; k < g.size(); ++k) c.push_back(to_id(g[k], first));
                conn.push_back(std::move(c));
                if (has_es) upsert_elset(es, {eid});
            }
        } else if (kw == "elset" || kw == "nset") {
            std::string nm;
            if (!param("elset", &nm) && !param("nset", &nm)) continue;
            const bool gen = params.count("generate") != 0;
            for (const Line* d : data) {
                std::vector<std::string> f = fields(d->text);
                std::vector<uint32_t> ids;
                if (gen) {
                    if (f.size() < 2) throw ParseFail{d->no, "GENERATE expects start, end[, step]"};
                    const int64_t a = to_int(f[0], d->no), b = to_int(f[1], d->no);
                    const int64_t step = f.size() == 3 ? to_int(f[2], d->no) : 1;
                    if (step == 0) throw ParseFail{d->no, "GENERATE step must not be zero"};
                    for (int64_t v = a; step > 0 ? v <= b : v >= b; v += step) {
                        if (v < 0 || v > (int64_t)UINT32_MAX) throw ParseFail{d->no, "id out of range"};
                        ids.push_back((uint32_t)v);
                    }
                } else {
                    for (const std::string& s : f) ids.push_back(to_id(s, d->no));
                }
                if (kw == "elset") {
                    upsert_elset(nm, ids);
                } else {
                    std::vector<uint32_t>& v = nsets[upper(nm)];
                    v.insert(v.end(), ids.begin(), ids.end());
                }
            }
        } else if (kw == "material") {
            std::string nm;
            if (param("name", &nm)) pending_material = upper(nm), have_pending = true;
        } else if (kw == "elastic") {
            if (!data.empty() && have_pending) {
                std::vector<std::string> f = fields(data[0]->text);
                if (f.size() < 2) throw ParseFail{data[0]->no, "*ELASTIC expects E, nu"};
                M.youngs.push_back(to_real(f[0], data[0]->no));
                M.poisson.push_back(to_real(f[1], data[0]->no));
                mat_names.push_back(pending_material);
                have_pending = false;
            }
        } else if (kw == "solid section") {
            std::string es, mat;
            if (param("elset", &es) && param("material", &mat) && !es.empty() && !mat.empty()) {
                assignments.emplace_back(upper(es), upper(mat));
                if (!data.empty()) {
                    std::vector<std::string> f = fields(data[0]->text);
                    if (f.empty()) throw ParseFail{data[0]->no, "*SOLID SECTION expects a thickness"};
                    M.thickness = to_real(f[0], data[0]->no);
                }
            }
        } else if (kw == "step") {
            if (cur) steps.push_back(*cur);
            std::string op;
            param("op", &op);
            const LoadSet& base = lower(op) == "new" ? LoadSet{} : (steps.empty() ? model_state : steps.back());
            cur = std::make_unique<LoadSet>(base);
        } else if (kw == "end step") {
            if (cur) {
                steps.push_back(*cur);
                cur.reset();
            }
        } else if (kw == "boundary" || kw == "cload") {
            LoadSet& target = cur ? *cur : model_state;
            std::string op;
            param("op", &op);
            const bool op_new = cur && lower(op) == "new";
            if (kw == "boundary") {
                if (op_new) target.bcs.clear();
                for (const Line* d : data) {
                    std::vector<std::string> f = fields(d->text);
                    if (f.size() < 2) throw ParseFail{d->no, "*BOUNDARY expects node/set, dof[, dof2]"};
                    std::vector<uint32_t> nodes = resolve(f[0], d->no);
                    const int64_t d1 = to_int(f[1], d->no);
                    std::vector<int32_t> dofs;
                    if (f.size() >= 3)
                        for (int64_t v = d1; v <= to_int(f[2], d->no); ++v) dofs.push_back((int32_t)v);
                    else
                        dofs.push_back((int32_t)d1);
                    target.bcs.emplace_back(std::move(nodes), std::move(dofs));
                }
            } else {
                if (op_new) target.cloads.clear();
                for (const Line* d : data) {
                    std::vector<std::string> f = fields(d->text);
                    if (f.size() != 3)
                        throw ParseFail{d->no, "*CLOAD expects node/set, dof, value; got " + std::to_string(f.size()) +
                                                   " field(s)"};
                    const int32_t dof = (int32_t)to_int(f[1], d->no);
                    const double val = to_real(f[2], d->no);
                    for (uint32_t nid : resolve(f[0], d->no)) target.cloads.emplace_back(nid, dof, val);
                }
            }
        }
    }
    if (cur) steps.push_back(*cur);
    const LoadSet& state = steps.empty() ? model_state : steps.back();

    for (const auto& c : conn) {
        M.conn_ids.insert(M.conn_ids.end(), c.begin(), c.end());
        M.conn_ptr.push_back((int64_t)M.conn_ids.size());
    }
    for (const auto& s : elsets) {
        M.elset_members.insert(M.elset_members.end(), s.second.begin(), s.second.end());
        M.elset_ptr.push_back((int64_t)M.elset_members.size());
    }
    // assignments by name: the first elset / material of that name, or -1
    for (const auto& a : assignments) {
        int32_t es = -1, mat = -1;
        for (size_t k = 0; k < elsets.size() && es < 0; ++k)
            if (elsets[k].first == a.first) es = (int32_t)k;
        for (size_t k = 0; k < mat_names.size() && mat < 0; ++k)
            if (mat_names[k] == a.second) mat = (int32_t)k;
        M.assign_elset.push_back(es);
        M.assign_material.push_back(mat);
    }
    for (const auto& b : state.bcs) {
        M.bc_nodes.insert(M.bc_nodes.end(), b.first.begin(), b.first.end());
        M.bc_node_ptr.push_back((int64_t)M.bc_nodes.size());
        M.bc_dofs.insert(M.bc_dofs.end(), b.second.begin(), b.second.end());
        M.bc_dof_ptr.push_back((int64_t)M.bc_dofs.size());
    }
    for (const auto& c : state.cloads) {
        M.cload_node.push_back(std::get<0>(c));
        M.cload_dof.push_back(std::get<1>(c));
        M.cload_value.push_back(std::get<2>(c));
    }
}

}  // namespace

int64_t parse_inp(const char* text, size_t len, OwnedModel& out, std::string* msg) {
    try {
        read(text, len, out);
        out.finalize();
        return 0;
    } catch (const ParseFail& f) {
        if (msg) *msg = f.msg;
        return f.line > 0 ? f.line : 1;
    }
}

}  // namespace svx

struct svx_deck {
    svx::OwnedModel model;
};

extern "C" {

int svx_deck_parse(const char* text, size_t len, svx_deck** out, svx_error* err) {
    if (err) {
        std::memset(err, 0, sizeof *err);
        err->index = -1;
    }
    if (!out || (!text && len)) return SVX_E_ARGUMENT;
    *out = nullptr;
    try {
        auto d = std::make_unique<svx_deck>();
        std::string msg;
        const int64_t line = svx::parse_inp(text, len, d->model, &msg);
        if (line) {
            if (err) {
                err->code = SVX_E_PARSE;
                err->index = line;
                snprintf(err->message, sizeof err->message, "line %lld: %s", (long long)line, msg.c_str());
            }
            return SVX_E_PARSE;
        }
        *out = d.release();
        return SVX_OK;
    } catch (const std::bad_alloc&) {
        return SVX_E_DEVICE;
    }
}

const svx_model_desc* svx_deck_model(const svx_deck* deck) { return deck ? &deck->model.desc : nullptr; }

void svx_deck_destroy(svx_deck* deck) { delete deck; }

}  // extern "C"
