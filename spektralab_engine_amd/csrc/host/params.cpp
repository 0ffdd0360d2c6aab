// The parameter schema, values and delta validation; rules stated in params.h.
#include "params.h"

#include <cmath>
#include <cstdio>

#include "util.h"

namespace spk::params {

namespace {

constexpr bool S = true, PR = false;  // layer: shoot / print
constexpr double kNo = 0.0;

constexpr Row num(const char* name, const char* path, Kind k, double lo, double hi, double def, bool shoot,
                  bool live = false) {
    return Row{name, path, k, true, lo, hi, def, nullptr, shoot, live};
}
constexpr Row flag(const char* name, const char* path, bool def, bool shoot) {
    return Row{name, path, Kind::Bool, false, kNo, kNo, def ? 1.0 : 0.0, nullptr, shoot, false};
}
constexpr Row str(const char* name, const char* path, const char* def, bool shoot) {
    return Row{name, path, Kind::Str, false, kNo, kNo, kNo, def, shoot, false};
}

// docs/API_SPEC.md section 3, row for row.
constexpr std::array<Row, kFields> kRows = {{
    str("film_stock", "film.info.stock", "kodak_portra_400", S),
    str("print_stock", "print.info.stock", "kodak_portra_endura", PR),
    num("exposure_compensation_ev", "camera.exposure_compensation_ev", Kind::Float, -8, 8, 0, S),
    flag("auto_exposure", "camera.auto_exposure", true, S),
    str("auto_exposure_method", "camera.auto_exposure_method", "center_weighted", S),
    num("film_format_mm", "camera.film_format_mm", Kind::Float, 4, 200, 35, S),
    num("lens_blur_um", "camera.lens_blur_um", Kind::Float, 0, 200, 0, S),
    flag("halation_active", "film_render.halation.active", true, S),
    num("halation_amount", "film_render.halation.halation_amount", Kind::Float, 0, 4, 1, S),
    num("halation_boost_ev", "film_render.halation.boost_ev", Kind::Float, 0, 8, 0, S),
    flag("grain_active", "film_render.grain.active", true, S),
    flag("grain_sublayers_active", "film_render.grain.sublayers_active", true, S),
    flag("dir_couplers_active", "film_render.dir_couplers.active", true, S),
    num("dir_couplers_amount", "film_render.dir_couplers.amount", Kind::Float, 0, 4, 1, S),
    num("density_curve_gamma", "film_render.density_curve_gamma", Kind::Float, 0.2, 4, 1, S),
    str("input_color_space", "io.input_color_space", "ProPhoto RGB", S),
    flag("input_cctf_decoding", "io.input_cctf_decoding", false, S),
    num("geometry_crop_x", "io.geometry.crop_x", Kind::Float, 0, 1, 0, S),
    num("geometry_crop_y", "io.geometry.crop_y", Kind::Float, 0, 1, 0, S),
    num("geometry_crop_w", "io.geometry.crop_w", Kind::Float, 1e-4, 1, 1, S),
    num("geometry_crop_h", "io.geometry.crop_h", Kind::Float, 1e-4, 1, 1, S),
    num("geometry_rotation_deg", "io.geometry.rotation_deg", Kind::Float, -45, 45, 0, S),
    num("geometry_quarter_turns", "io.geometry.quarter_turns", Kind::Int, 0, 3, 0, S),
    flag("geometry_flip_h", "io.geometry.flip_h", false, S),
    flag("geometry_flip_v", "io.geometry.flip_v", false, S),
    num("print_exposure", "enlarger.print_exposure", Kind::Float, 0.05, 20, 1, PR, true),
    num("m_filter_shift", "enlarger.m_filter_shift", Kind::Float, -1, 1, 0, PR, true),
    num("y_filter_shift", "enlarger.y_filter_shift", Kind::Float, -1, 1, 0, PR, true),
    num("c_filter_neutral", "enlarger.c_filter_neutral", Kind::Float, 0, 200, 0, PR),
    num("m_filter_neutral", "enlarger.m_filter_neutral", Kind::Float, 0, 200, 65, PR),
    num("y_filter_neutral", "enlarger.y_filter_neutral", Kind::Float, 0, 200, 55, PR),
    num("preflash_exposure", "enlarger.preflash_exposure", Kind::Float, 0, 1, 0, PR, true),
    str("enlarger_illuminant", "enlarger.illuminant", "TH-KG3", PR),
    flag("glare_active", "print_render.glare.active", true, PR),
    flag("scanner_white_correction", "scanner.white_correction", false, PR),
    flag("scanner_black_correction", "scanner.black_correction", false, PR),
    num("scanner_lens_blur", "scanner.lens_blur", Kind::Float, 0, 20, 0, PR),
    str("output_color_space", "io.output_color_space", "ProPhoto RGB", PR),
    flag("output_cctf_encoding", "io.output_cctf_encoding", true, PR),
    flag("scan_film", "io.scan_film", false, PR),
    flag("extended_dynamic_range", "print_render.edr_enabled", false, PR),
    num("halation_scatter_amount", "film_render.halation.scatter_amount", Kind::Float, 0, 1, 1, S),
    num("grain_amount", "film_render.grain.amount", Kind::Float, 0, 2, 1, S),
    num("glare_amount", "print_render.glare.amount", Kind::Float, 0, 30, 1, PR),
    flag("antihalation_removed", "film_render.halation.antihalation_removed", false, S),
    num("filter_shift_scale", "enlarger.filter_shift_scale", Kind::Float, 1, 60, 1, PR, true),
    num("halation_headroom_ev", "film_render.halation.headroom_ev", Kind::Float, 0, 8, 4, S),
    num("preview_long_edge", "io.preview_long_edge", Kind::Int, 800, 8192, 1600, PR),
}};

// The accepted values of auto_exposure_method, in the order its message lists them.
constexpr std::array<const char*, 7> kMeterMethods = {"center_weighted", "average",          "median",
                                                      "balanced",        "center",           "protect_highlights",
                                                      "protect_shadows"};

const char* kind_word(Kind k) {
    switch (k) {
        case Kind::Float: return "float";
        case Kind::Int: return "int";
        case Kind::Bool: return "bool";
        case Kind::Str: return "str";
    }
    return "?";
}

std::string fixed(double v) {  // printf's %f, the form of the range message
    char buf[512];
    std::snprintf(buf, sizeof buf, "%f", v);
    return buf;
}

json::Value default_json(const Row& r) {
    switch (r.kind) {
        case Kind::Str: return json::Value(r.text);
        case Kind::Bool: return json::Value(r.number != 0.0);
        default: return json::Value(r.number);
    }
}

// Checks one delta entry against its row; returns the value to store.
Slot checked(const Row& r, const json::Value& v) {
    const std::string name = r.name;
    const bool kind_ok = r.kind == Kind::Str ? v.is_text() : r.kind == Kind::Bool ? v.is_bool() : v.is_number();
    if (!kind_ok) fail_user("'" + name + "' expects " + kind_word(r.kind));
    Slot out;
    if (r.kind == Kind::Str) {
        out.str = v.text();
        if (std::string_view(r.name) == "auto_exposure_method") {
            bool known = false;
            std::string list;
            for (const char* m : kMeterMethods) {
                known = known || out.str == m;
                list += list.empty() ? m : std::string(", ") + m;
            }
            if (!known) fail_user("'auto_exposure_method' = '" + out.str + "' is not one of " + list);
        }
        return out;
    }
    if (r.kind == Kind::Bool) {
        out.num = v.boolean() ? 1.0 : 0.0;
        return out;
    }
    const double x = v.number();
    if (std::isnan(x)) fail_user("'" + name + "' expects " + kind_word(r.kind) + ", not NaN");
    if (r.ranged && !(x >= r.lo && x <= r.hi))
        fail_user("'" + name + "'=" + fixed(x) + " is outside [" + fixed(r.lo) + ", " + fixed(r.hi) + "]");
    out.num = r.kind == Kind::Int ? std::round(x) : x;  // ints: half away from zero, after the range check
    return out;
}

}  // namespace

const std::array<Row, kFields>& rows() { return kRows; }

std::optional<P> lookup(std::string_view name) {
    for (size_t i = 0; i < kFields; ++i)
        if (name == kRows[i].name) return P(i);
    return std::nullopt;
}

const std::string& schema_json() {
    static const std::string text = [] {
        json::Value fields = json::Value::list();
        for (const Row& r : kRows) {
            json::Value f = json::Value::map();
            f.set("default", default_json(r));
            f.set("layer", r.shoot ? "shoot" : "print");
            f.set("live", r.live);
            f.set("name", r.name);
            f.set("path", r.path);
            if (r.ranged) f.set("range", json::Value::list({r.lo, r.hi}));
            f.set("type", kind_word(r.kind));
            fields.append(std::move(f));
        }
        json::Value doc = json::Value::map();
        doc.set("fields", std::move(fields));
        doc.set("schema_version", 1);
        return json::dump(doc);
    }();
    return text;
}

Values::Values() {
    for (size_t i = 0; i < kFields; ++i) {
        slots_[i].num = kRows[i].number;
        if (kRows[i].text) slots_[i].str = kRows[i].text;
    }
}

json::Value Values::to_json() const {
    json::Value o = json::Value::map();
    for (size_t i = 0; i < kFields; ++i) {
        const Row& r = kRows[i];
        if (r.kind == Kind::Str) o.set(r.name, slots_[i].str);
        else if (r.kind == Kind::Bool) o.set(r.name, slots_[i].num != 0.0);
        else o.set(r.name, slots_[i].num);
    }
    return o;
}

Delta parse_delta(std::string_view text, Origin origin) {
    Delta d;
    if (text.empty()) {
        if (origin == Origin::Open) return d;
        fail_user("params_delta must be an object");
    }
    json::Value doc;
    try {
        doc = json::parse(text);
    } catch (const json::SyntaxError& e) {
        fail_user(e.what());
    }
    if (!doc.is_map()) fail_user("params_delta must be an object");
    for (const json::Member& m : doc.members()) {  // ascending bytewise by key
        if (origin == Origin::Open && m.first == "product_defaults") {
            if (!m.second.is_bool()) fail_user("'product_defaults' expects bool");
            d.product_defaults = m.second.boolean();
            continue;
        }
        const std::optional<P> p = lookup(m.first);
        if (!p) fail_user("unknown parameter '" + m.first + "'");
        const Row& r = row(*p);
        d.sets.emplace_back(*p, checked(r, m.second));
        if (*p == P::film_stock || *p == P::print_stock) d.names_stock = true;
        if (r.shoot) d.touches_shoot = true;
        if (!r.live) d.needs_rebuild = true;
    }
    return d;
}

void apply(Values& v, const Delta& d) {
    for (const auto& [p, s] : d.sets) v.put(p, s);
}

}  // namespace spk::params
