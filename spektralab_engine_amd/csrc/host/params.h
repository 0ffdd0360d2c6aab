// Render parameters: the 48-field schema of the ABI, one session's values, and parameter deltas.
//
// The schema (names, dotted paths, types, ranges, defaults, layers, live flags, in params_schema
// order) is the data contract of reference docs/API_SPEC.md section 3; delta validation follows
// section 2.4 of the same document.
#pragma once

#include <array>
#include <cstdint>
#include <optional>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

#include "json.h"

namespace spk::params {

// One identifier per schema field, in params_schema order (the field's own name).
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
};
inline constexpr size_t kFields = size_t(P::preview_long_edge) + 1;

enum class Kind : uint8_t { Float, Int, Bool, Str };

// One schema row.  A ranged field has lo <= hi; `text` is the default of a Str field, `number`
// the default of the others (bools as 0 / 1).  Print-layer fields have shoot = false.
struct Row {
    const char* name;
    const char* path;
    Kind kind;
    bool ranged;
    double lo, hi;
    double number;
    const char* text;
    bool shoot;
    bool live;
};

// The rows in params_schema order; row(p) is the row of field p.
const std::array<Row, kFields>& rows();
inline const Row& row(P p) { return rows()[size_t(p)]; }
std::optional<P> lookup(std::string_view name);
// {"fields": [...], "schema_version": 1} (the params_schema reply).
const std::string& schema_json();

// A field's value: numbers and bools in `num` (an int is a whole number, a bool 0 or 1), strings
// in `str`.
struct Slot {
    double num = 0.0;
    std::string str;
};

// The values of every field (schema defaults until set).
class Values {
public:
    Values();
    double f(P p) const { return slots_[size_t(p)].num; }
    int64_t i(P p) const { return int64_t(slots_[size_t(p)].num); }
    bool b(P p) const { return slots_[size_t(p)].num != 0.0; }
    const std::string& s(P p) const { return slots_[size_t(p)].str; }
    const Slot& slot(P p) const { return slots_[size_t(p)]; }
    void put(P p, Slot v) { slots_[size_t(p)] = std::move(v); }
    void put_number(P p, double v) { slots_[size_t(p)].num = v; }
    void put_text(P p, std::string v) { slots_[size_t(p)].str = std::move(v); }
    // The flat params object of the replies: 48 keys, numbers, bools and strings.
    json::Value to_json() const;

private:
    std::array<Slot, kFields> slots_;
};
using ParamSet = Values;

// A validated delta: the fields it sets (ascending by name) and what they touch.
struct Delta {
    std::vector<std::pair<P, Slot>> sets;
    bool names_stock = false;   // film_stock or print_stock
    bool touches_shoot = false; // some field of the shoot layer
    bool needs_rebuild = false; // some field that is neither live nor a stock
    std::optional<bool> product_defaults;
};

// Where a delta comes from: an open delta may carry "product_defaults" (bool, not a field).
enum class Origin { Open, SetParams };
// Parses and validates a delta text (an empty open delta is {}); failures throw SPK_ERR_USER with
// the message section 2.4 names.
Delta parse_delta(std::string_view text, Origin origin);
void apply(Values& v, const Delta& d);

}  // namespace spk::params
