// Schema rows, ParamSet and delta validation; rules stated in params.h.
#include "params.h"

#include <cmath>
#include <cstdio>

#include "util.h"

namespace spk::params {

namespace {

using R = std::optional<std::pair<double, double>>;

FieldDecl row(const char* name, const char* path, Type t, R range, json::Value def, Layer l, bool live = false) {
    FieldDecl f;
    f.name = name;
    f.path = path;
    f.type = t;
    f.range = range;
    f.def = std::move(def);
    f.layer = l;
    f.live = live;
    return f;
}

const char* type_name(Type t) {
    switch (t) {
        case Type::Float: return "float";
        case Type::Int: return "int";
        case Type::Bool: return "bool";
        case Type::Str: return "str";
    }
    return "?";
}

std::vector<FieldDecl> make_rows() {
    const R none = std::nullopt;
    auto rg = [](double lo, double hi) { return R(std::make_pair(lo, hi)); };
    constexpr Layer S = Layer::Shoot, Pr = Layer::Print;
    std::vector<FieldDecl> v = {
        row("film_stock", "film.info.stock", Type::Str, none, "kodak_portra_400", S),
        row("print_stock", "print.info.stock", Type::Str, none, "kodak_portra_endura", Pr),
        row("exposure_compensation_ev", "camera.exposure_compensation_ev", Type::Float, rg(-8, 8), 0, S),
        row("auto_exposure", "camera.auto_exposure", Type::Bool, none, true, S),
        row("auto_exposure_method", "camera.auto_exposure_method", Type::Str, none, "center_weighted", S),
        row("film_format_mm", "camera.film_format_mm", Type::Float, rg(4, 200), 35, S),
        row("lens_blur_um", "camera.lens_blur_um", Type::Float, rg(0, 200), 0, S),
        row("halation_active", "film_render.halation.active", Type::Bool, none, true, S),
        row("halation_amount", "film_render.halation.halation_amount", Type::Float, rg(0, 4), 1, S),
        row("halation_boost_ev", "film_render.halation.boost_ev", Type::Float, rg(0, 8), 0, S),
        row("grain_active", "film_render.grain.active", Type::Bool, none, true, S),
        row("grain_sublayers_active", "film_render.grain.sublayers_active", Type::Bool, none, true, S),
        row("dir_couplers_active", "film_render.dir_couplers.active", Type::Bool, none, true, S),
        row("dir_couplers_amount", "film_render.dir_couplers.amount", Type::Float, rg(0, 4), 1, S),
        row("density_curve_gamma", "film_render.density_curve_gamma", Type::Float, rg(0.2, 4), 1, S),
        row("input_color_space", "io.input_color_space", Type::Str, none, "ProPhoto RGB", S),
        row("input_cctf_decoding", "io.input_cctf_decoding", Type::Bool, none, false, S),
        row("geometry_crop_x", "io.geometry.crop_x", Type::Float, rg(0, 1), 0, S),
        row("geometry_crop_y", "io.geometry.crop_y", Type::Float, rg(0, 1), 0, S),
        row("geometry_crop_w", "io.geometry.crop_w", Type::Float, rg(1e-4, 1), 1, S),
        row("geometry_crop_h", "io.geometry.crop_h", Type::Float, rg(1e-4, 1), 1, S),
        row("geometry_rotation_deg", "io.geometry.rotation_deg", Type::Float, rg(-45, 45), 0, S),
        row("geometry_quarter_turns", "io.geometry.quarter_turns", Type::Int, rg(0, 3), 0, S),
        row("geometry_flip_h", "io.geometry.flip_h", Type::Bool, none, false, S),
        row("geometry_flip_v", "io.geometry.flip_v", Type::Bool, none, false, S),
        row("print_exposure", "enlarger.print_exposure", Type::Float, rg(0.05, 20), 1, Pr, true),
        row("m_filter_shift", "enlarger.m_filter_shift", Type::Float, rg(-1, 1), 0, Pr, true),
        row("y_filter_shift", "enlarger.y_filter_shift", Type::Float, rg(-1, 1), 0, Pr, true),
        row("c_filter_neutral", "enlarger.c_filter_neutral", Type::Float, rg(0, 200), 0, Pr),
        row("m_filter_neutral", "enlarger.m_filter_neutral", Type::Float, rg(0, 200), 65, Pr),
        row("y_filter_neutral", "enlarger.y_filter_neutral", Type::Float, rg(0, 200), 55, Pr),
        row("preflash_exposure", "enlarger.preflash_exposure", Type::Float, rg(0, 1), 0, Pr, true),
        row("enlarger_illuminant", "enlarger.illuminant", Type::Str, none, "TH-KG3", Pr),
        row("glare_active", "print_render.glare.active", Type::Bool, none, true, Pr),
        row("scanner_white_correction", "scanner.white_correction", Type::Bool, none, false, Pr),
        row("scanner_black_correction", "scanner.black_correction", Type::Bool, none, false, Pr),
        row("scanner_lens_blur", "scanner.lens_blur", Type::Float, rg(0, 20), 0, Pr),
        row("output_color_space", "io.output_color_space", Type::Str, none, "ProPhoto RGB", Pr),
        row("output_cctf_encoding", "io.output_cctf_encoding", Type::Bool, none, true, Pr),
        row("scan_film", "io.scan_film", Type::Bool, none, false, Pr),
        row("extended_dynamic_range", "print_render.edr_enabled", Type::Bool, none, false, Pr),
        row("halation_scatter_amount", "film_render.halation.scatter_amount", Type::Float, rg(0, 1), 1, S),
        row("grain_amount", "film_render.grain.amount", Type::Float, rg(0, 2), 1, S),
        row("glare_amount", "print_render.glare.amount", Type::Float, rg(0, 30), 1, Pr),
        row("antihalation_removed", "film_render.halation.antihalation_removed", Type::Bool, none, false, S),
        row("filter_shift_scale", "enlarger.filter_shift_scale", Type::Float, rg(1, 60), 1, Pr, true),
        row("halation_headroom_ev", "film_render.halation.headroom_ev", Type::Float, rg(0, 8), 4, S),
        row("preview_long_edge", "io.preview_long_edge", Type::Int, rg(800, 8192), 1600, Pr),
    };
    static const std::vector<std::string> methods = {"center_weighted", "average", "median", "balanced",
                                                     "center", "protect_highlights", "protect_shadows"};
    v[id(P::auto_exposure_method)].check_enum = [](std::string_view s) -> std::optional<std::string> {
        std::string list;
        for (const std::string& m : methods) {
            if (m == s) return std::nullopt;
            list += (list.empty() ? "" : ", ") + m;
        }
        return "'auto_exposure_method' = '" + std::string(s) + "' is not one of " + list;
    };
    return v;
}

std::string fmt_f(double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%f", v);
    return buf;
}

bool type_matches(Type t, const json::Value& v) {
    switch (t) {
        case Type::Float:
        case Type::Int: return v.is_number();
        case Type::Bool: return v.is_bool();
        case Type::Str: return v.is_string();
    }
    return false;
}

}  // namespace

std::shared_ptr<const Schema> Schema::build() {
    static const std::vector<FieldDecl> rows = make_rows();
    auto s = std::make_shared<Schema>();
    s->fields_ = rows;
    json::Array arr;
    for (const FieldDecl& f : s->fields_) {
        json::Value r;
        r["name"] = f.name;
        r["path"] = f.path;
        r["type"] = type_name(f.type);
        r["layer"] = f.layer == Layer::Shoot ? "shoot" : "print";
        r["default"] = f.def;
        r["live"] = f.live;
        if (f.range) r["range"] = json::Array{f.range->first, f.range->second};
        arr.push_back(std::move(r));
    }
    json::Value doc;
    doc["schema_version"] = 1;
    doc["fields"] = std::move(arr);
    s->json_ = json::write(doc);
    return s;
}

std::optional<size_t> Schema::find(std::string_view name) const {
    for (size_t i = 0; i < fields_.size(); ++i)
        if (fields_[i].name == name) return i;
    return std::nullopt;
}

ParamSet::ParamSet(std::shared_ptr<const Schema> s) : schema_(std::move(s)) {
    for (size_t i = 0; i < schema_->size(); ++i) values_.push_back(schema_->field(i).def);
}

json::Value ParamSet::to_json() const {
    json::Value o = json::Object{};
    for (size_t i = 0; i < values_.size(); ++i) o[schema_->field(i).name] = values_[i];
    return o;
}

Delta validate_delta(const Schema& schema, std::string_view text, DeltaContext ctx) {
    Delta d;
    if (text.empty() && ctx == DeltaContext::Open) return d;
    json::Value doc;
    try {
        doc = json::parse(text);
    } catch (const json::ParseError& e) {
        fail_user(e.what());
    }
    if (!doc.is_object()) fail_user("params_delta must be an object");
    json::Object obj = doc.as_object();
    if (ctx == DeltaContext::Open) {
        if (auto it = obj.find("product_defaults"); it != obj.end()) {
            if (!it->second.is_bool()) fail_user("'product_defaults' expects bool");
            d.product_defaults = it->second.as_bool();
            obj.erase(it);
        }
    }
    for (const auto& [name, value] : obj) {  // std::map: ascending bytewise order
        const std::optional<size_t> fid = schema.find(name);
        if (!fid) fail_user("unknown parameter '" + name + "'");
        const FieldDecl& f = schema.field(*fid);
        if (!type_matches(f.type, value)) fail_user("'" + name + "' expects " + type_name(f.type));
        json::Value stored = value;
        if (f.type == Type::Str && f.check_enum)
            if (auto msg = f.check_enum(value.as_string())) fail_user(*msg);
        if (f.type == Type::Float || f.type == Type::Int) {
            const double v = value.as_number();
            if (f.range) {
                const auto [lo, hi] = *f.range;
                if (!(v >= lo && v <= hi))
                    fail_user("'" + name + "'=" + fmt_f(v) + " is outside [" + fmt_f(lo) + ", " + fmt_f(hi) + "]");
            } else if (std::isnan(v)) {
                fail_user("'" + name + "' expects " + type_name(f.type) + ", not NaN");
            }
            if (f.type == Type::Int) stored = json::Value(std::round(v));
        }
        d.fields.emplace_back(*fid, std::move(stored));
        if (name == "film_stock" || name == "print_stock") d.has_stock = true;
        if (f.layer == Layer::Shoot) d.any_shoot = true;
        if (!f.live) d.any_rebuild = true;
    }
    if (d.has_stock) d.any_rebuild = true;
    return d;
}

void apply_delta(ParamSet& p, const Delta& d) {
    for (const auto& [fid, v] : d.fields) p.set(fid, v);
}

}  // namespace spk::params
