// The parameter schema (48 rows), the flat parameter values and delta validation.
//
// Rows, types, ranges, defaults, layers and live flags restate the reference's
// src/params/schema.cpp:51-113; validation restates src/params/delta.cpp:40-96 (keys in ascending
// bytewise order, first failure wins, NaN refused, ints rounded half away from zero).
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

#include "json.h"

namespace spk::params {

enum class Type : uint8_t { Float, Int, Bool, Str };
enum class Layer : uint8_t { Shoot, Print };

struct FieldDecl {
    std::string name, path;
    Type type = Type::Float;
    Layer layer = Layer::Shoot;
    bool live = false;
    std::optional<std::pair<double, double>> range;
    json::Value def;
    std::function<std::optional<std::string>(std::string_view)> check_enum;
};

enum class P : uint16_t {
    film_stock, print_stock,
    exposure_compensation_ev, auto_exposure, auto_exposure_method, film_format_mm, lens_blur_um,
    halation_active, halation_amount, halation_boost_ev,
    grain_active, grain_sublayers_active,
    dir_couplers_active, dir_couplers_amount, density_curve_gamma,
    input_color_space, input_cctf_decoding,
    geometry_crop_x, geometry_crop_y, geometry_crop_w, geometry_crop_h, geometry_rotation_deg,
    geometry_quarter_turns, geometry_flip_h, geometry_flip_v,
    print_exposure, m_filter_shift, y_filter_shift, c_filter_neutral, m_filter_neutral, y_filter_neutral,
    preflash_exposure, enlarger_illuminant,
    glare_active, scanner_white_correction, scanner_black_correction, scanner_lens_blur,
    output_color_space, output_cctf_encoding, scan_film, extended_dynamic_range,
    halation_scatter_amount, grain_amount, glare_amount, antihalation_removed, filter_shift_scale,
    halation_headroom_ev, preview_long_edge,
    kCount
};
inline constexpr size_t id(P p) { return size_t(p); }

class Schema {
public:
    static std::shared_ptr<const Schema> build();
    size_t size() const { return fields_.size(); }
    const FieldDecl& field(size_t i) const { return fields_[i]; }
    std::optional<size_t> find(std::string_view name) const;
    const std::string& json_text() const { return json_; }

private:
    std::vector<FieldDecl> fields_;
    std::string json_;
};

class ParamSet {
public:
    explicit ParamSet(std::shared_ptr<const Schema> s);
    double f(P p) const { return values_[id(p)].as_number(); }
    int64_t i(P p) const { return int64_t(values_[id(p)].as_number()); }
    bool b(P p) const { return values_[id(p)].as_bool(); }
    const std::string& s(P p) const { return values_[id(p)].as_string(); }
    const json::Value& get(size_t fid) const { return values_[fid]; }
    void set(size_t fid, json::Value v) { values_[fid] = std::move(v); }
    void set(P p, json::Value v) { set(id(p), std::move(v)); }
    json::Value to_json() const;
    const Schema& schema() const { return *schema_; }

private:
    std::shared_ptr<const Schema> schema_;
    std::vector<json::Value> values_;
};

struct Delta {
    std::vector<std::pair<size_t, json::Value>> fields;
    bool has_stock = false, any_shoot = false, any_rebuild = false;
    std::optional<bool> product_defaults;
};

enum class DeltaContext { Open, SetParams };
Delta validate_delta(const Schema& schema, std::string_view text, DeltaContext ctx);
void apply_delta(ParamSet& p, const Delta& d);

}  // namespace spk::params
