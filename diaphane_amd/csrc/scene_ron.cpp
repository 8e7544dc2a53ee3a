// Scene files: the reference's RON text format (src/scene.rs:467-498, where
// serde derives it from the types; grid.rs:376-392 and 889-935 for the grid
// recipe and its checks). A recursive-descent reader of the RON subset those
// types produce, and a writer whose output the reader takes back exactly.
//
// What the reader accepts, following serde's derived rules over ron:
//  * structs as `Name(field: value, ...)` with the name optional, fields in
//    any order, a trailing comma allowed; a missing field is an error unless
//    serde gives it a default (`velocity`, `refinements`, `max_ratio`); an
//    unknown field is skipped, as serde does without deny_unknown_fields;
//  * enums as `Variant` or `Variant(field: value, ...)`;
//  * `[f32; 3]` as a tuple `(a, b, c)` (a list `[a, b, c]` is taken too);
//  * numbers with sign, fraction and exponent, `inf`, `-inf`, `NaN`; f32
//    fields are rounded once from the decimal text (strtof), as Rust's f32
//    parser does;
//  * `//` line comments and nested `/* */` block comments.
// Errors name the line and column, and the grid's checks (the GridSpec →
// Grid conversion) come back as errors with the reference's wording.
#include <algorithm>
#include <cerrno>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <set>
#include <string>

#include "scene.hpp"

namespace dpa {

namespace {

class Reader {
public:
    explicit Reader(const std::string &text) : s_(text) {}

    [[noreturn]] void error(const std::string &what) const { error_at(i_, what); }

    [[noreturn]] void error_at(size_t at, const std::string &what) const {
        size_t line = 1, column = 1;
        for (size_t k = 0; k < at && k < s_.size(); ++k) {
            if (s_[k] == '\n') {
                ++line;
                column = 1;
            } else {
                ++column;
            }
        }
        throw Error{DP_ERR_INVALID, std::to_string(line) + ":" + std::to_string(column) + ": " + what};
    }

    void skip_space() {
        for (;;) {
            while (i_ < s_.size() && std::strchr(" \t\r\n", s_[i_])) ++i_;
            if (starts("//")) {
                while (i_ < s_.size() && s_[i_] != '\n') ++i_;
            } else if (starts("/*")) {
                const size_t open = i_;
                int depth = 0;
                do {
                    if (starts("/*")) {
                        ++depth;
                        i_ += 2;
                    } else if (starts("*/")) {
                        --depth;
                        i_ += 2;
                    } else if (i_ < s_.size()) {
                        ++i_;
                    } else {
                        error_at(open, "unclosed block comment");
                    }
                } while (depth > 0);
            } else {
                return;
            }
        }
    }

    bool at_end() {
        skip_space();
        return i_ >= s_.size();
    }

    char peek() {
        skip_space();
        return i_ < s_.size() ? s_[i_] : '\0';
    }

    bool accept(char c) {
        if (peek() != c) return false;
        ++i_;
        return true;
    }

    void expect(char c) {
        if (!accept(c)) error(std::string("expected '") + c + "'" + found());
    }

    std::string found() {
        skip_space();
        if (i_ >= s_.size()) return ", found the end of the text";
        return std::string(", found '") + s_[i_] + "'";
    }

    bool ident_start() {
        const char c = peek();
        return std::isalpha(static_cast<unsigned char>(c)) || c == '_';
    }

    std::string ident() {
        if (!ident_start()) error("expected an identifier" + found());
        const size_t begin = i_;
        while (i_ < s_.size() && (std::isalnum(static_cast<unsigned char>(s_[i_])) || s_[i_] == '_')) ++i_;
        return s_.substr(begin, i_ - begin);
    }

    size_t position() {
        skip_space();
        return i_;
    }

    // A number token: sign, digits, fraction, exponent, or inf / NaN.
    std::string number_token() {
        skip_space();
        const size_t begin = i_;
        if (i_ < s_.size() && (s_[i_] == '+' || s_[i_] == '-')) ++i_;
        if (s_.compare(i_, 3, "inf") == 0 || s_.compare(i_, 3, "NaN") == 0) {
            i_ += 3;
        } else {
            while (i_ < s_.size() && (std::isdigit(static_cast<unsigned char>(s_[i_])) || s_[i_] == '.' ||
                                      s_[i_] == '_'))
                ++i_;
            if (i_ < s_.size() && (s_[i_] == 'e' || s_[i_] == 'E')) {
                ++i_;
                if (i_ < s_.size() && (s_[i_] == '+' || s_[i_] == '-')) ++i_;
                while (i_ < s_.size() && std::isdigit(static_cast<unsigned char>(s_[i_]))) ++i_;
            }
        }
        std::string token = s_.substr(begin, i_ - begin);
        token.erase(std::remove(token.begin(), token.end(), '_'), token.end());
        if (token.empty() || token == "+" || token == "-") error_at(begin, "expected a number" + found());
        return token;
    }

    float f32() {
        const size_t at = position();
        const std::string token = number_token();
        if (token == "inf" || token == "+inf") return INFINITY;
        if (token == "-inf") return -INFINITY;
        if (token == "NaN" || token == "+NaN" || token == "-NaN") return NAN;
        char *end = nullptr;
        errno = 0;
        const float value = std::strtof(token.c_str(), &end);
        if (end != token.c_str() + token.size()) error_at(at, "'" + token + "' is not a number");
        return value;  // overflow rounds to ±inf, as Rust's parser does
    }

    uint32_t u32() {
        const size_t at = position();
        const std::string token = number_token();
        if (token.find_first_not_of("0123456789+") != std::string::npos || token[0] == '+')
            error_at(at, "expected an unsigned integer, found '" + token + "'");
        char *end = nullptr;
        errno = 0;
        const unsigned long long value = std::strtoull(token.c_str(), &end, 10);
        if (errno == ERANGE || value > 0xFFFFFFFFull) error_at(at, "'" + token + "' does not fit in u32");
        return static_cast<uint32_t>(value);
    }

    // `(a, b, c)` or `[a, b, c]`.
    void triple(float out[3]) {
        const char close = accept('(') ? ')' : (accept('[') ? ']' : '\0');
        if (!close) error("expected a 3-tuple" + found());
        for (int k = 0; k < 3; ++k) {
            out[k] = f32();
            if (k < 2) expect(',');
        }
        accept(',');
        expect(close);
    }

    // The body of a struct or struct variant: `(name: value, ...)`. `field`
    // reads the value of a known name and returns false for an unknown one,
    // which is then skipped. Returns the set of names read.
    std::set<std::string> fields(const std::string &type, const std::function<bool(const std::string &)> &field) {
        std::set<std::string> seen;
        expect('(');
        while (!accept(')')) {
            const size_t at = position();
            const std::string name = ident();
            expect(':');
            if (!seen.insert(name).second) error_at(at, "duplicate field `" + name + "` in " + type);
            if (!field(name)) skip_value();
            if (!accept(',')) {
                expect(')');
                break;
            }
        }
        return seen;
    }

    // An optional struct name, checked against the expected one.
    void struct_name(const std::string &expected) {
        if (!ident_start()) return;
        const size_t at = position();
        const std::string name = ident();
        if (name != expected) error_at(at, "expected " + expected + ", found `" + name + "`");
    }

    template <typename T>
    void list(const std::function<T()> &item, std::vector<T> &out) {
        expect('[');
        while (!accept(']')) {
            out.push_back(item());
            if (!accept(',')) {
                expect(']');
                break;
            }
        }
    }

    // Any value, for an unknown field.
    void skip_value() {
        const char c = peek();
        if (c == '(' || c == '[' || c == '{') {
            skip_group();
        } else if (c == '"') {
            ++i_;
            while (i_ < s_.size() && s_[i_] != '"') i_ += s_[i_] == '\\' ? 2 : 1;
            expect('"');
        } else if (ident_start() && s_.compare(i_, 3, "inf") != 0 && s_.compare(i_, 3, "NaN") != 0) {
            ident();
            if (peek() == '(') skip_group();
        } else {
            number_token();
        }
    }

    void skip_group() {
        const size_t open = position();
        int depth = 0;
        do {
            const char c = peek();
            if (c == '\0') error_at(open, "unclosed bracket");
            if (c == '(' || c == '[' || c == '{') ++depth;
            if (c == ')' || c == ']' || c == '}') --depth;
            if (c == '"') {
                skip_value();
                continue;
            }
            ++i_;
        } while (depth > 0);
    }

private:
    bool starts(const char *prefix) const { return s_.compare(i_, std::strlen(prefix), prefix) == 0; }

    const std::string &s_;
    size_t i_ = 0;
};

void require(Reader &r, size_t at, const std::set<std::string> &seen, const std::string &type,
             std::initializer_list<const char *> names) {
    for (const char *name : names)
        if (!seen.count(name)) r.error_at(at, std::string("missing field `") + name + "` in " + type);
}

uint32_t axis(Reader &r) {
    const size_t at = r.position();
    const std::string name = r.ident();
    if (name == "X") return 0;
    if (name == "Y") return 1;
    if (name == "Z") return 2;
    r.error_at(at, "unknown axis `" + name + "`, expected X, Y or Z");
}

void grid(Reader &r, SceneFile &file) {
    const size_t at = r.position();
    r.struct_name("Grid");
    dp_grid_spec &g = file.scene.grid;
    g.max_ratio = 1.15f;  // src/grid.rs:401-403
    const auto seen = r.fields("Grid", [&](const std::string &name) {
        if (name == "extent") {
            r.struct_name("Extent");
            const auto inner = r.fields("Extent", [&](const std::string &n) {
                if (n == "x") g.extent[0] = r.u32();
                else if (n == "y") g.extent[1] = r.u32();
                else if (n == "z") g.extent[2] = r.u32();
                else return false;
                return true;
            });
            require(r, at, inner, "Extent", {"x", "y", "z"});
        } else if (name == "cell_size") {
            g.cell_size = r.f32();
        } else if (name == "courant") {
            g.courant = r.f32();
        } else if (name == "max_ratio") {
            g.max_ratio = r.f32();
        } else if (name == "refinements") {
            r.list<dp_refinement>(
                [&] {
                    const size_t rat = r.position();
                    r.struct_name("Refinement");
                    dp_refinement f{};
                    const auto got = r.fields("Refinement", [&](const std::string &n) {
                        if (n == "center") r.triple(f.center);
                        else if (n == "size") r.triple(f.size);
                        else if (n == "cell_size") f.cell_size = r.f32();
                        else return false;
                        return true;
                    });
                    require(r, rat, got, "Refinement", {"center", "size", "cell_size"});
                    return f;
                },
                file.refinements);
        } else {
            return false;
        }
        return true;
    });
    require(r, at, seen, "Grid", {"extent", "cell_size", "courant"});
    // The GridSpec → Grid conversion's checks (src/grid.rs:894-935).
    char buffer[256];
    if (g.extent[0] < 2 || g.extent[1] < 2 || g.extent[2] < 2) {
        std::snprintf(buffer, sizeof buffer,
                      "a grid needs at least 2 cells on every axis, got Extent { x: %u, y: %u, z: %u }", g.extent[0],
                      g.extent[1], g.extent[2]);
        r.error_at(at, buffer);
    }
    if (!(g.cell_size > 0.0f && std::isfinite(g.cell_size))) {
        std::snprintf(buffer, sizeof buffer, "cell_size must be positive and finite, got %g", double(g.cell_size));
        r.error_at(at, buffer);
    }
    if (!(g.courant > 0.0f && g.courant <= kCourantLimit)) {
        std::snprintf(buffer, sizeof buffer, "courant must lie in (0, %.8g], got %g", double(kCourantLimit),
                      double(g.courant));
        r.error_at(at, buffer);
    }
    if (!(g.max_ratio > 1.0f && std::isfinite(g.max_ratio))) {
        std::snprintf(buffer, sizeof buffer, "max_ratio must exceed 1, got %g", double(g.max_ratio));
        r.error_at(at, buffer);
    }
    for (const dp_refinement &f : file.refinements)
        if (!(f.cell_size > 0.0f && std::isfinite(f.cell_size))) {
            std::snprintf(buffer, sizeof buffer, "a refinement's cell_size must be positive and finite, got %g",
                          double(f.cell_size));
            r.error_at(at, buffer);
        }
}

void boundary(Reader &r, dp_scene &scene) {
    const size_t at = r.position();
    const std::string variant = r.ident();
    if (variant == "Pec") {
        scene.boundary = DP_BOUNDARY_PEC;
        return;
    }
    if (variant != "Absorbing") r.error_at(at, "unknown boundary `" + variant + "`, expected Pec or Absorbing");
    scene.boundary = DP_BOUNDARY_ABSORBING;
    const auto seen = r.fields("Absorbing", [&](const std::string &name) {
        if (name == "thickness") scene.absorber_thickness = r.u32();
        else if (name == "target_reflection") scene.target_reflection = r.f32();
        else return false;
        return true;
    });
    require(r, at, seen, "Absorbing", {"thickness", "target_reflection"});
}

void materials(Reader &r, SceneFile &file) {
    const size_t at = r.position();
    r.struct_name("MaterialTable");
    const auto seen = r.fields("MaterialTable", [&](const std::string &name) {
        if (name != "materials") return false;
        r.list<dp_material>(
            [&] {
                const size_t mat = r.position();
                r.struct_name("Material");
                dp_material m{};
                const auto got = r.fields("Material", [&](const std::string &n) {
                    if (n == "relative_permittivity") m.relative_permittivity = r.f32();
                    else if (n == "relative_permeability") m.relative_permeability = r.f32();
                    else if (n == "conductivity") m.conductivity = r.f32();
                    else if (n == "magnetic_loss") m.magnetic_loss = r.f32();
                    else return false;
                    return true;
                });
                require(r, mat, got, "Material",
                        {"relative_permittivity", "relative_permeability", "conductivity", "magnetic_loss"});
                return m;
            },
            file.materials);
        return true;
    });
    require(r, at, seen, "MaterialTable", {"materials"});
}

dp_shape shape(Reader &r) {
    const size_t at = r.position();
    const std::string variant = r.ident();
    dp_shape s{};
    if (variant == "Block") {
        s.kind = DP_SHAPE_BLOCK;
        const auto seen = r.fields("Block", [&](const std::string &n) {
            if (n == "center") r.triple(s.center);
            else if (n == "size") r.triple(s.size);
            else if (n == "material") s.material = r.u32();
            else return false;
            return true;
        });
        require(r, at, seen, "Block", {"center", "size", "material"});
    } else if (variant == "Sphere") {
        s.kind = DP_SHAPE_SPHERE;
        const auto seen = r.fields("Sphere", [&](const std::string &n) {
            if (n == "center") r.triple(s.center);
            else if (n == "radius") s.radius = r.f32();
            else if (n == "material") s.material = r.u32();
            else return false;
            return true;
        });
        require(r, at, seen, "Sphere", {"center", "radius", "material"});
    } else if (variant == "Slab") {
        s.kind = DP_SHAPE_SLAB;
        const auto seen = r.fields("Slab", [&](const std::string &n) {
            if (n == "axis") s.axis = axis(r);
            else if (n == "offset") s.offset = r.f32();
            else if (n == "thickness") s.thickness = r.f32();
            else if (n == "material") s.material = r.u32();
            else return false;
            return true;
        });
        require(r, at, seen, "Slab", {"axis", "offset", "thickness", "material"});
    } else {
        r.error_at(at, "unknown shape `" + variant + "`, expected Block, Sphere or Slab");
    }
    return s;
}

void waveform(Reader &r, dp_waveform &w) {
    const size_t at = r.position();
    const std::string variant = r.ident();
    if (variant == "Ricker") {
        w.kind = DP_WAVE_RICKER;
        const auto seen = r.fields("Ricker", [&](const std::string &n) {
            if (n == "peak_frequency") w.frequency = r.f32();
            else if (n == "delay") w.delay = r.f32();
            else return false;
            return true;
        });
        require(r, at, seen, "Ricker", {"peak_frequency", "delay"});
    } else if (variant == "GaussianPulse") {
        w.kind = DP_WAVE_GAUSSIAN_PULSE;
        const auto seen = r.fields("GaussianPulse", [&](const std::string &n) {
            if (n == "center_frequency") w.frequency = r.f32();
            else if (n == "duration_cycles") w.cycles = r.f32();
            else if (n == "delay") w.delay = r.f32();
            else return false;
            return true;
        });
        require(r, at, seen, "GaussianPulse", {"center_frequency", "duration_cycles", "delay"});
    } else if (variant == "ContinuousWave") {
        w.kind = DP_WAVE_CONTINUOUS_WAVE;
        const auto seen = r.fields("ContinuousWave", [&](const std::string &n) {
            if (n == "frequency") w.frequency = r.f32();
            else if (n == "ramp_cycles") w.cycles = r.f32();
            else return false;
            return true;
        });
        require(r, at, seen, "ContinuousWave", {"frequency", "ramp_cycles"});
    } else if (variant == "DifferentiatedGaussian") {  // this engine's addition
        w.kind = DP_WAVE_DIFF_GAUSSIAN;
        const auto seen = r.fields("DifferentiatedGaussian", [&](const std::string &n) {
            if (n == "width") w.width = r.f32();
            else if (n == "delay") w.delay = r.f32();
            else return false;
            return true;
        });
        require(r, at, seen, "DifferentiatedGaussian", {"width", "delay"});
    } else {
        r.error_at(at, "unknown waveform `" + variant + "`, expected Ricker, GaussianPulse or ContinuousWave");
    }
}

dp_source source(Reader &r) {
    const size_t at = r.position();
    r.struct_name("Source");
    dp_source s{};
    const auto seen = r.fields("Source", [&](const std::string &name) {
        if (name == "shape") {
            const size_t sat = r.position();
            const std::string variant = r.ident();
            if (variant == "Point") {
                s.shape = DP_SOURCE_POINT;
                const auto got = r.fields("Point", [&](const std::string &n) {
                    if (n == "at") r.triple(s.at);
                    else if (n == "velocity") r.triple(s.velocity);  // serde default: zero
                    else return false;
                    return true;
                });
                require(r, sat, got, "Point", {"at"});
            } else if (variant == "Sheet") {
                s.shape = DP_SOURCE_SHEET;
                const auto got = r.fields("Sheet", [&](const std::string &n) {
                    if (n == "axis") s.axis = axis(r);
                    else if (n == "offset") s.offset = r.f32();
                    else if (n == "waist") s.waist = r.f32();
                    else return false;
                    return true;
                });
                require(r, sat, got, "Sheet", {"axis", "offset", "waist"});
            } else if (variant == "PlaneWave") {
                s.shape = DP_SOURCE_PLANE_WAVE;
                const auto got = r.fields("PlaneWave", [&](const std::string &n) {
                    if (n == "axis") s.axis = axis(r);
                    else if (n == "offset") s.offset = r.f32();
                    else return false;
                    return true;
                });
                require(r, sat, got, "PlaneWave", {"axis", "offset"});
            } else {
                r.error_at(sat, "unknown source shape `" + variant + "`, expected Point, Sheet or PlaneWave");
            }
        } else if (name == "polarization") {
            s.polarization = axis(r);
        } else if (name == "waveform") {
            waveform(r, s.waveform);
        } else if (name == "amplitude") {
            s.amplitude = r.f32();
        } else {
            return false;
        }
        return true;
    });
    require(r, at, seen, "Source", {"shape", "polarization", "waveform", "amplitude"});
    return s;
}

// Shortest decimal that reads back as the same f32 (at most 9 digits).
std::string f32_text(float v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buffer[48];
    for (int digits = 1; digits <= 9; ++digits) {
        std::snprintf(buffer, sizeof buffer, "%.*g", digits, double(v));
        if (std::strtof(buffer, nullptr) == v) break;
    }
    std::string text = buffer;
    // RON reads a bare integer as an integer; keep floats visibly floats.
    if (text.find_first_of(".eEn") == std::string::npos) text += ".0";
    return text;
}

std::string triple_text(const float v[3]) {
    return "(" + f32_text(v[0]) + ", " + f32_text(v[1]) + ", " + f32_text(v[2]) + ")";
}

const char *axis_text(uint32_t a) { return a == 0 ? "X" : (a == 1 ? "Y" : "Z"); }

}  // namespace

std::unique_ptr<SceneFile> parse_scene_ron(const std::string &text) {
    auto file = std::make_unique<SceneFile>();
    Reader r(text);
    const size_t at = r.position();
    r.struct_name("Scene");
    dp_scene &scene = file->scene;
    const auto seen = r.fields("Scene", [&](const std::string &name) {
        if (name == "grid") grid(r, *file);
        else if (name == "boundary") boundary(r, scene);
        else if (name == "materials") materials(r, *file);
        else if (name == "shapes") r.list<dp_shape>([&] { return shape(r); }, file->shapes);
        else if (name == "sources") r.list<dp_source>([&] { return source(r); }, file->sources);
        else return false;
        return true;
    });
    require(r, at, seen, "Scene", {"grid", "boundary", "materials", "shapes", "sources"});
    if (!r.at_end()) r.error("trailing characters after the scene");
    file->bind();
    return file;
}

void SceneFile::bind() {
    scene.grid.refinements = refinements.data();
    scene.grid.refinement_count = static_cast<uint32_t>(refinements.size());
    scene.materials = materials.data();
    scene.material_count = static_cast<uint32_t>(materials.size());
    scene.shapes = shapes.data();
    scene.shape_count = static_cast<uint32_t>(shapes.size());
    scene.sources = sources.data();
    scene.source_count = static_cast<uint32_t>(sources.size());
}

std::string scene_to_ron(const dp_scene &scene) {
    std::string o;
    const dp_grid_spec &g = scene.grid;
    o += "Scene(\n    grid: Grid(\n";
    o += "        extent: Extent(x: " + std::to_string(g.extent[0]) + ", y: " + std::to_string(g.extent[1]) +
         ", z: " + std::to_string(g.extent[2]) + "),\n";
    o += "        cell_size: " + f32_text(g.cell_size) + ",\n";
    o += "        courant: " + f32_text(g.courant) + ",\n";
    o += "        refinements: [\n";
    for (uint32_t k = 0; k < g.refinement_count; ++k) {
        const dp_refinement &f = g.refinements[k];
        o += "            Refinement(center: " + triple_text(f.center) + ", size: " + triple_text(f.size) +
             ", cell_size: " + f32_text(f.cell_size) + "),\n";
    }
    o += "        ],\n        max_ratio: " + f32_text(g.max_ratio) + ",\n    ),\n";
    if (scene.boundary == DP_BOUNDARY_PEC)
        o += "    boundary: Pec,\n";
    else
        o += "    boundary: Absorbing(thickness: " + std::to_string(scene.absorber_thickness) +
             ", target_reflection: " + f32_text(scene.target_reflection) + "),\n";
    o += "    materials: MaterialTable(materials: [\n";
    for (uint32_t k = 0; k < scene.material_count; ++k) {
        const dp_material &m = scene.materials[k];
        o += "        Material(relative_permittivity: " + f32_text(m.relative_permittivity) +
             ", relative_permeability: " + f32_text(m.relative_permeability) +
             ", conductivity: " + f32_text(m.conductivity) + ", magnetic_loss: " + f32_text(m.magnetic_loss) +
             "),\n";
    }
    o += "    ]),\n    shapes: [\n";
    for (uint32_t k = 0; k < scene.shape_count; ++k) {
        const dp_shape &s = scene.shapes[k];
        const std::string material = ", material: " + std::to_string(s.material) + "),\n";
        if (s.kind == DP_SHAPE_BLOCK)
            o += "        Block(center: " + triple_text(s.center) + ", size: " + triple_text(s.size) + material;
        else if (s.kind == DP_SHAPE_SPHERE)
            o += "        Sphere(center: " + triple_text(s.center) + ", radius: " + f32_text(s.radius) + material;
        else
            o += std::string("        Slab(axis: ") + axis_text(s.axis) + ", offset: " + f32_text(s.offset) +
                 ", thickness: " + f32_text(s.thickness) + material;
    }
    o += "    ],\n    sources: [\n";
    for (uint32_t k = 0; k < scene.source_count; ++k) {
        const dp_source &s = scene.sources[k];
        o += "        Source(\n            shape: ";
        if (s.shape == DP_SOURCE_POINT)
            o += "Point(at: " + triple_text(s.at) + ", velocity: " + triple_text(s.velocity) + ")";
        else if (s.shape == DP_SOURCE_SHEET)
            o += std::string("Sheet(axis: ") + axis_text(s.axis) + ", offset: " + f32_text(s.offset) +
                 ", waist: " + f32_text(s.waist) + ")";
        else
            o += std::string("PlaneWave(axis: ") + axis_text(s.axis) + ", offset: " + f32_text(s.offset) + ")";
        o += std::string(",\n            polarization: ") + axis_text(s.polarization) + ",\n            waveform: ";
        const dp_waveform &w = s.waveform;
        if (w.kind == DP_WAVE_RICKER)
            o += "Ricker(peak_frequency: " + f32_text(w.frequency) + ", delay: " + f32_text(w.delay) + ")";
        else if (w.kind == DP_WAVE_GAUSSIAN_PULSE)
            o += "GaussianPulse(center_frequency: " + f32_text(w.frequency) +
                 ", duration_cycles: " + f32_text(w.cycles) + ", delay: " + f32_text(w.delay) + ")";
        else if (w.kind == DP_WAVE_CONTINUOUS_WAVE)
            o += "ContinuousWave(frequency: " + f32_text(w.frequency) + ", ramp_cycles: " + f32_text(w.cycles) +
                 ")";
        else
            o += "DifferentiatedGaussian(width: " + f32_text(w.width) + ", delay: " + f32_text(w.delay) + ")";
        o += ",\n            amplitude: " + f32_text(s.amplitude) + ",\n        ),\n";
    }
    o += "    ],\n)\n";
    return o;
}

}  // namespace dpa
