// lbm_hip/io.hpp -- the reference's front-end formats without boost,
// pugixml or VTK: the config file and command line
// (include/io/configuration.h:14-161), the scenario XML
// (include/io/scenario.h:21-188, grammar in build/scenarios/README.xml), the
// legacy STRUCTURED_POINTS fluid mask (include/io/vtk.hpp:94-157) and a .vts
// writer of density and velocity (include/io/vtk.hpp:15-89).
#pragma once

#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbm_hip/lbm.hpp"

namespace lbm {
namespace io {

// ------------------------------------------------------------------ Config --
// Options of the "Configuration options" group (configuration.h:70-86), read
// from `key = value` lines of the positional input file and from
// --key value / --key=value on the command line; command-line values win,
// unknown command-line options are ignored. Two extensions: `lattice`
// (D3Q15/D3Q19/D3Q27, default D3Q19) and `storage` (f64/f32/f16/bf16).
class Config {
    std::string _collision_model{"bgk"};
    std::string _input_file;
    std::string _output_dir{"vtk"};
    std::string _output_filename{"output"};
    std::string _scenario_xml;
    uint64_t _timesteps{0};
    uint64_t _timesteps_per_plot{0};
    double _tau{1.0};
    uint32_t _omp_threads{1};
    std::string _lattice{"D3Q19"};
    std::string _storage{"f64"};
    std::string _layout{"ab"};

    static std::string trim(const std::string& s) {
        size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
        return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    }
    static const std::map<std::string, std::string>& aliases() {
        static const std::map<std::string, std::string> m = {
            {"c", "collision-model"}, {"t", "timesteps"}, {"i", "input-file"}, {"h", "help"}};
        return m;
    }
    static bool known(const std::string& k) {
        static const char* keys[] = {"collision-model", "tau", "timesteps", "timesteps-per-plot",
                                     "scenario-file", "omp-threads", "output-dir", "lattice",
                                     "storage", "layout"};
        for (const char* s : keys)
            if (k == s) return true;
        return false;
    }
    void apply(const std::string& k, const std::string& v) {
        try {
            if (k == "collision-model") _collision_model = v;
            else if (k == "tau") _tau = std::stod(v);
            else if (k == "timesteps") _timesteps = std::stoull(v);
            else if (k == "timesteps-per-plot") _timesteps_per_plot = std::stoull(v);
            else if (k == "scenario-file") _scenario_xml = v;
            else if (k == "omp-threads") _omp_threads = static_cast<uint32_t>(std::stoul(v));
            else if (k == "output-dir") _output_dir = v;
            else if (k == "lattice") _lattice = v;
            else if (k == "storage") _storage = v;
            else if (k == "layout") _layout = v;
        } catch (const std::exception&) {
            throw std::logic_error("the argument ('" + v + "') for option '--" + k + "' is invalid");
        }
    }

public:
    bool help = false;

    Config() = default;
    // configuration.h:68-145 (the output directory is emptied as there).
    Config(int argc, char** argv, bool touch_output_dir = true) {
        std::map<std::string, std::string> cli;
        std::vector<std::string> positional;
        for (int i = 1; i < argc; ++i) {
            std::string a = argv[i];
            if (a.size() > 1 && a[0] == '-') {
                std::string k = a.substr(a[1] == '-' ? 2 : 1), v;
                const size_t eq = k.find('=');
                if (eq != std::string::npos) {
                    v = k.substr(eq + 1);
                    k = k.substr(0, eq);
                } else if (k != "help" && k != "h" && i + 1 < argc) {
                    v = argv[++i];
                }
                auto al = aliases().find(k);
                if (al != aliases().end()) k = al->second;
                if (k == "help") help = true;
                else if (k == "input-file") positional.push_back(v);
                else if (known(k)) cli.emplace(k, v);  // first occurrence wins
            } else {
                positional.push_back(a);
            }
        }
        std::cout << "LBM simulation (MI355X collide-stream step)" << std::endl;
        if (help || argc == 1) {
            help = true;
            std::cout << "Usage: lbm [file] additional-options...\n"
                      << "  --collision-model bgk --tau T --timesteps N --timesteps-per-plot P\n"
                      << "  --scenario-file F --omp-threads N --output-dir D\n"
                      << "  --lattice D3Q15|D3Q19|D3Q27 --storage f64|f32|f16|bf16|f64s --layout ab|aa" << std::endl;
            std::exit(1);  // as configuration.h:109-114
        }
        std::map<std::string, std::string> file;
        if (!positional.empty()) {
            _input_file = positional.front();
            std::cout << "Reading configuration file..." << std::endl;
            std::ifstream in(_input_file);
            if (!in) throw std::logic_error("can not read options configuration file '" + _input_file + "'");
            std::string line;
            while (std::getline(in, line)) {
                const size_t hash = line.find('#');
                if (hash != std::string::npos) line = line.substr(0, hash);
                line = trim(line);
                if (line.empty()) continue;
                const size_t eq = line.find('=');
                if (eq == std::string::npos) throw std::logic_error("invalid line: " + line);
                const std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
                if (!known(k)) throw std::logic_error("unrecognised option '" + k + "'");
                file.emplace(k, v);
            }
        }
        for (auto& kv : file)
            if (!cli.count(kv.first)) cli.insert(kv);
        for (const char* req : {"tau", "timesteps", "timesteps-per-plot", "scenario-file"})
            if (!cli.count(req))
                throw std::logic_error(std::string("the option '--") + req + "' is required but missing");
        for (auto& kv : cli) apply(kv.first, kv.second);
        // configuration.h:123-128 (checked always, not only in debug builds)
        if (_timesteps == 0) throw std::logic_error("timesteps must be > 0");
        if (!(_tau > 0.5 && _tau < 2.0)) throw std::logic_error("tau must be in (0.5, 2.0)");
        if (_omp_threads == 0) throw std::logic_error("omp-threads must be > 0");
        if (_collision_model != "bgk") throw std::logic_error("only the bgk collision model is supported");
        if (touch_output_dir) {
            namespace fs = std::filesystem;
            fs::path dir(_output_dir);
            if (!fs::exists(dir)) {
                std::cout << "Output directory \"" + _output_dir + "\" does not exist. Creating." << std::endl;
                fs::create_directories(dir);
            } else {
                for (auto& e : fs::directory_iterator(dir))
                    if (e.is_regular_file()) fs::remove(e.path());
            }
        }
    }

    std::string collision_model() const { return _collision_model; }
    std::string input_file() const { return _input_file; }
    std::string output_dir() const { return _output_dir; }
    std::string output_filename() const { return _output_filename; }
    void set_output_filename(const std::string& s) { _output_filename = s; }
    uint64_t timesteps() const { return _timesteps; }
    uint64_t timesteps_per_plot() const { return _timesteps_per_plot; }
    double tau() const { return _tau; }
    uint32_t omp_threads() const { return _omp_threads; }
    std::string scenario_xml() const { return _scenario_xml; }
    std::string lattice() const { return _lattice; }
    std::string storage() const { return _storage; }
    int layout_id() const {
        if (_layout == "ab") return LBM_LAYOUT_AB;
        if (_layout == "aa") return LBM_LAYOUT_AA;
        throw std::logic_error("layout must be ab or aa");
    }
    int storage_id() const {
        if (_storage == "f64") return LBM_STORAGE_F64;
        if (_storage == "f32") return LBM_STORAGE_F32;
        if (_storage == "f16") return LBM_STORAGE_F16;
        if (_storage == "bf16") return LBM_STORAGE_BF16;
        if (_storage == "f64s") return LBM_STORAGE_F64S;
        throw std::logic_error("storage must be f64, f32, f16, bf16 or f64s");
    }
};

inline std::ostream& operator<<(std::ostream& o, const Config& c) {  // configuration.h:148-161
    o << "Configuration:" << '\n'
      << "> Configuration file:     " << c.input_file() << '\n'
      << "> Output directory:       " << c.output_dir() << '\n'
      << "> Output filename:        " << c.output_filename() << '\n'
      << "> Collision model:        " << c.collision_model() << '\n'
      << "> Lattice / storage:      " << c.lattice() << " / " << c.storage() << '\n'
      << "> Tau:                    " << c.tau() << '\n'
      << "> Timesteps:              " << c.timesteps() << '\n'
      << "> Timesteps per plot:     " << c.timesteps_per_plot() << '\n'
      << "> Scenario file:          " << c.scenario_xml();
    return o;
}

// --------------------------------------------------------------- mini XML --
// Elements with attributes; text, comments, declarations and doctypes skipped.
struct XmlNode {
    std::string name;
    std::map<std::string, std::string> attrs;
    std::vector<XmlNode> children;
    const XmlNode* child(const std::string& n) const {
        for (auto& c : children)
            if (c.name == n) return &c;
        return nullptr;
    }
    bool has(const std::string& a) const { return attrs.count(a) != 0; }
    const std::string& attr(const std::string& a) const {
        auto it = attrs.find(a);
        if (it == attrs.end())  // scenario.h:21-26
            throw std::logic_error("Missing attribute \"" + a + "\" for node \"" + name + "\"!");
        return it->second;
    }
};

inline XmlNode parse_xml(const std::string& text) {
    size_t i = 0;
    auto fail = [&](const std::string& m) { throw std::logic_error("XML: " + m + " at offset " + std::to_string(i)); };
    auto skip_ws = [&] { while (i < text.size() && std::isspace(static_cast<unsigned char>(text[i]))) ++i; };
    XmlNode root;
    root.name = "#document";
    std::vector<XmlNode*> stack{&root};
    while (i < text.size()) {
        const size_t lt = text.find('<', i);
        if (lt == std::string::npos) break;
        i = lt;
        if (text.compare(i, 4, "<!--") == 0) {
            const size_t e = text.find("-->", i + 4);
            if (e == std::string::npos) fail("unterminated comment");
            i = e + 3;
        } else if (text.compare(i, 2, "<?") == 0 || text.compare(i, 2, "<!") == 0) {
            const size_t e = text.find('>', i);
            if (e == std::string::npos) fail("unterminated declaration");
            i = e + 1;
        } else if (text.compare(i, 2, "</") == 0) {
            const size_t e = text.find('>', i);
            if (e == std::string::npos) fail("unterminated end tag");
            std::string n = text.substr(i + 2, e - i - 2);
            while (!n.empty() && std::isspace(static_cast<unsigned char>(n.back()))) n.pop_back();
            if (stack.size() < 2 || stack.back()->name != n) fail("mismatched </" + n + ">");
            stack.pop_back();
            i = e + 1;
        } else {
            ++i;
            XmlNode node;
            while (i < text.size() && !std::isspace(static_cast<unsigned char>(text[i])) && text[i] != '>' && text[i] != '/')
                node.name += text[i++];
            bool self_close = false;
            for (;;) {
                skip_ws();
                if (i >= text.size()) fail("unterminated tag");
                if (text[i] == '/') {
                    self_close = true;
                    ++i;
                    continue;
                }
                if (text[i] == '>') {
                    ++i;
                    break;
                }
                std::string an;
                while (i < text.size() && text[i] != '=' && !std::isspace(static_cast<unsigned char>(text[i]))) an += text[i++];
                skip_ws();
                if (i >= text.size() || text[i] != '=') fail("attribute without value");
                ++i;
                skip_ws();
                if (i >= text.size() || (text[i] != '"' && text[i] != '\'')) fail("unquoted attribute");
                const char qch = text[i++];
                const size_t e = text.find(qch, i);
                if (e == std::string::npos) fail("unterminated attribute");
                node.attrs[an] = text.substr(i, e - i);
                i = e + 1;
            }
            stack.back()->children.push_back(node);
            if (!self_close) stack.push_back(&stack.back()->children.back());
        }
    }
    if (stack.size() != 1) fail("unclosed element <" + stack.back()->name + ">");
    return root;
}

// -------------------------------------------------------- VTK fluid mask --
struct VtkMask {
    size_t xl = 0, yl = 0, zl = 0;
    double origin[3] = {0, 0, 0};
    double spacing[3] = {1, 1, 1};
    std::vector<uint8_t> fluid;  // xl*yl*zl, x fastest, 1 = fluid
};

// Legacy ASCII STRUCTURED_POINTS (include/io/vtk.hpp:94-157).
inline VtkMask read_vtk_mask(const std::string& filename) {
    std::ifstream in(filename);
    if (!in)
        throw std::logic_error("VTK file \"" + filename + "\" does not exist or does not seem to be a valid structured grids file!");
    VtkMask m;
    std::string line, word;
    bool dims = false, data = false, ascii = false, sp = false;
    while (!data && std::getline(in, line)) {
        std::istringstream ls(line);
        if (!(ls >> word)) continue;
        for (auto& ch : word) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
        if (word == "ASCII") ascii = true;
        else if (word == "DATASET") {
            ls >> word;
            sp = (word == "STRUCTURED_POINTS");
        } else if (word == "DIMENSIONS") {
            ls >> m.xl >> m.yl >> m.zl;
            dims = true;
        } else if (word == "ORIGIN") ls >> m.origin[0] >> m.origin[1] >> m.origin[2];
        else if (word == "SPACING" || word == "ASPECT_RATIO") ls >> m.spacing[0] >> m.spacing[1] >> m.spacing[2];
        else if (word == "LOOKUP_TABLE") data = true;
    }
    if (!ascii || !sp || !dims || !data)
        throw std::logic_error("VTK file \"" + filename + "\" is not an ASCII STRUCTURED_POINTS mask");
    const size_t n = m.xl * m.yl * m.zl;
    m.fluid.resize(n);
    for (size_t i = 0; i < n; ++i) {
        long v;
        if (!(in >> v)) throw std::logic_error("VTK file \"" + filename + "\": point data is short");
        m.fluid[i] = v != 0;
    }
    return m;
}

// -------------------------------------------------------------- scenario --
template <class M>
NonFluidCollision<M>& parse_condition(const XmlNode& b, Domain<M>& domain) {  // scenario.h:28-88
    using BF = BoundaryKeeper<M>;
    const std::string condition = b.attr("condition");
    auto num = [&](const char* a) { return std::stod(b.attr(a)); };
    if (condition == "noslip") return BF::template get_collision<NoSlipBoundary<M>>(domain);
    if (condition == "movingwall")
        return BF::template get_collision<MovingWallBoundary<M>>(domain, double_array<3>{num("vx"), num("vy"), num("vz")});
    if (condition == "freeslip") return BF::template get_collision<FreeSlipBoundary<M>>(domain);
    if (condition == "outflow")
        return BF::template get_collision<OutflowBoundary<M>>(domain, b.has("rho-ref") ? num("rho-ref") : 1.0);
    if (condition == "inflow")
        return BF::template get_collision<InflowBoundary<M>>(
            domain, double_array<3>{num("vx"), num("vy"), num("vz")}, b.has("rho-ref") ? num("rho-ref") : 1.0);
    if (condition == "pressure") return BF::template get_collision<PressureBoundary<M>>(domain, num("rho-in"));
    throw std::logic_error(condition + " boundary condition not supported!");
}

template <class M>
void parse_boundary(const XmlNode& b, Domain<M>& domain) {  // scenario.h:90-128
    const std::string extent = b.attr("extent");
    auto& cond = parse_condition<M>(b, domain);
    const size_t xl = domain.xlength(), yl = domain.ylength(), zl = domain.zlength();
    if (extent == "z0") domain.setBoundaryCondition(cond, 0, xl + 1, 0, yl + 1, 0, 0);
    else if (extent == "zmax") domain.setBoundaryCondition(cond, 0, xl + 1, 0, yl + 1, zl + 1, zl + 1);
    else if (extent == "x0") domain.setBoundaryCondition(cond, 0, 0, 0, yl + 1, 0, zl + 1);
    else if (extent == "xmax") domain.setBoundaryCondition(cond, xl + 1, xl + 1, 0, yl + 1, 0, zl + 1);
    else if (extent == "y0") domain.setBoundaryCondition(cond, 0, xl + 1, 0, 0, 0, zl + 1);
    else if (extent == "ymax") domain.setBoundaryCondition(cond, 0, xl + 1, yl + 1, yl + 1, 0, zl + 1);
    else {
        std::vector<uint64_t> ex;
        std::istringstream s(extent);
        std::string tok;
        while (s >> tok) {
            if (tok.find_first_not_of("0123456789") != std::string::npos)
                throw std::logic_error("bad lexical cast: \"" + tok + "\" in extent \"" + extent + "\"");
            ex.push_back(std::stoull(tok));
        }
        if (ex.size() < 6)
            throw std::logic_error("Extent \"" + extent + "\" is not complete! Must be six values!");
        domain.setBoundaryCondition(cond, ex[0], ex[1], ex[2], ex[3], ex[4], ex[5]);
    }
}

// scenario.h:131-188. Relative vtk-file paths are taken relative to the
// working directory, as in the reference.
template <class M>
Domain_ptr<M> parse_scenario_file(const std::string& filename, Config& cfg,
                                  FluidCollision<M>& collision, const DeviceOptions& opt) {
    std::ifstream in(filename);
    if (!in) throw std::logic_error("XML file \"" + filename + "\" could not be read properly!");
    std::stringstream ss;
    ss << in.rdbuf();
    XmlNode doc;
    try {
        doc = parse_xml(ss.str());
    } catch (const std::logic_error&) {
        throw std::logic_error("XML file \"" + filename + "\" could not be read properly!");
    }
    const XmlNode* scenario = doc.child("scenario");
    if (!scenario) throw std::logic_error("Scenario node missing!");
    if (!scenario->has("name")) throw std::logic_error("Scenario name is missing!");
    const std::string name = scenario->attrs.at("name");
    std::cout << "Reading scenario: \"" << name << "\"..." << std::endl;
    cfg.set_output_filename(name);
    const XmlNode* xd = scenario->child("domain");
    if (!xd) throw std::logic_error("Domain node is missing!");
    Domain_ptr<M> domain;
    if (xd->has("vtk-file")) {
        const VtkMask m = read_vtk_mask(xd->attrs.at("vtk-file"));
        domain = std::make_unique<Domain<M>>(m.xl, m.yl, m.zl, collision, m.origin[0], m.origin[1],
                                             m.origin[2], m.spacing[0], m.spacing[1], m.spacing[2], opt);
        domain->set_solid_mask(m.fluid);
    } else if (xd->has("xl") && xd->has("yl") && xd->has("zl")) {
        const long xl = std::stol(xd->attrs.at("xl")), yl = std::stol(xd->attrs.at("yl")),
                   zl = std::stol(xd->attrs.at("zl"));
        if (xl <= 0 || yl <= 0 || zl <= 0) throw std::logic_error("domain lengths must be positive");
        domain = std::make_unique<Domain<M>>(xl, yl, zl, collision, 0, 0, 0, 1, 1, 1, opt);
    } else {
        throw std::logic_error("Neither vtk-file nor xl/yl/zl attribute provided to domain node!");
    }
    for (const auto& b : xd->children)
        if (b.name == "boundary") parse_boundary(b, *domain);
    return domain;
}

// The reference's three-argument form: the storage mode comes from the
// configuration's `storage` key (fp64 in the reference's arithmetic when unset).
template <class M>
Domain_ptr<M> parse_scenario_file(const std::string& filename, Config& cfg,
                                  FluidCollision<M>& collision) {
    DeviceOptions opt;
    opt.storage = cfg.storage_id();
    opt.layout = cfg.layout_id();
    return parse_scenario_file<M>(filename, cfg, collision, opt);
}

// ---------------------------------------------------------------- output --
// vtk.hpp:41-89: points (xs*x+xo-1, ...), Velocity and Density of every
// interior cell from the post-collision field, file <dir>/<name>.<t>.vts.
// Written as VTK XML StructuredGrid with ASCII data arrays.
template <class M>
void write_vtk_file(const Domain<M>& domain, const std::string& output_dir,
                    const std::string& output_filename, uint64_t t) {
    const size_t xl = domain.xlength(), yl = domain.ylength(), zl = domain.zlength();
    std::vector<double> rho, u;
    domain.moments(rho, u);
    std::ostringstream path;
    path << output_dir << "/" << output_filename << "." << t << ".vts";
    FILE* f = std::fopen(path.str().c_str(), "w");
    if (!f) throw std::logic_error("cannot write " + path.str());
    std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    std::fprintf(f, "  <StructuredGrid WholeExtent=\"0 %zu 0 %zu 0 %zu\">\n", xl - 1, yl - 1, zl - 1);
    std::fprintf(f, "    <Piece Extent=\"0 %zu 0 %zu 0 %zu\">\n", xl - 1, yl - 1, zl - 1);
    std::fprintf(f, "      <PointData Scalars=\"Density\" Vectors=\"Velocity\">\n");
    std::fprintf(f, "        <DataArray type=\"Float64\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    for (size_t i = 0; i < rho.size(); ++i) std::fprintf(f, "%.17g %.17g %.17g\n", u[3 * i], u[3 * i + 1], u[3 * i + 2]);
    std::fprintf(f, "        </DataArray>\n        <DataArray type=\"Float64\" Name=\"Density\" format=\"ascii\">\n");
    for (size_t i = 0; i < rho.size(); ++i) std::fprintf(f, "%.17g\n", rho[i]);
    std::fprintf(f, "        </DataArray>\n      </PointData>\n      <Points>\n");
    std::fprintf(f, "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    for (size_t z = 1; z <= zl; ++z)
        for (size_t y = 1; y <= yl; ++y)
            for (size_t x = 1; x <= xl; ++x)
                std::fprintf(f, "%.17g %.17g %.17g\n", domain.xspacing() * x + domain.xorigin() - 1,
                             domain.yspacing() * y + domain.yorigin() - 1,
                             domain.zspacing() * z + domain.zorigin() - 1);
    std::fprintf(f, "        </DataArray>\n      </Points>\n    </Piece>\n  </StructuredGrid>\n</VTKFile>\n");
    std::fclose(f);
}

}  // namespace io
}  // namespace lbm
