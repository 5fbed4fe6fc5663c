// scene_file.cpp -- RISE scene files (native "RISE ASCII SCENE 7" text)
// read into the spectral path tracer's scene tables (host code; see
// include/rise_hip_scene_file.h for the supported chunks).
//
// Each chunk handler cites the reference chunk parser whose parameters and
// defaults it follows (paths relative to the reference's src/Library/).
#include "rise_hip_scene_file.h"

#include "api_common.h"
#include "bvh_build.h"
#include "pt_internal.h"
#include "risemesh_internal.h"
#include "rise_hip_rgb_uplift.h"
#include "rise_hip_output.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

using namespace rise_hip;

struct RiseMeshDeleter {
    void operator()(rise_hip_risemesh* m) const { rise_hip_risemesh_release(m); }
};
using RiseMeshPtr = std::shared_ptr<rise_hip_risemesh>;

struct rise_hip_scene_file {
    std::vector<std::vector<float>> meshes;   // world-space triangles per object (9 floats each)
    std::vector<RiseMeshPtr> object_risemesh; // per object: the .risemesh its geometry came from, if any
    std::vector<rise_hip_object_desc> objects;
    std::vector<rise_hip_object_media> object_media;
    std::vector<rise_hip_material_desc> materials;
    std::vector<std::vector<float>> amps;
    std::vector<rise_hip_spd_desc> spds;
    std::vector<rise_hip_medium_desc> media;
    bool has_sky = false;
    rise_hip_sky_desc sky{};
    std::vector<rise_hip_directional_light_desc> dir_lights;
    rise_hip_render_desc desc{};
    int skipped = 0;
    bool any_media = false;
    std::vector<rise_hip_output_desc> outputs;
    std::vector<std::string> output_patterns;
};

namespace {

constexpr double kDegToRad = 3.14159265358979323846 / 180.0;

struct Chunk {
    std::string keyword;
    int line = 0;
    std::vector<std::pair<std::string, std::string>> params;  // in file order (keys may repeat)

    const std::string* get(const std::string& k) const
    {
        const std::string* v = nullptr;
        for (const auto& p : params)
            if (p.first == k) v = &p.second;  // the last occurrence wins
        return v;
    }
    bool has(const std::string& k) const { return get(k) != nullptr; }
    std::vector<std::string> all(const std::string& k) const
    {
        std::vector<std::string> out;
        for (const auto& p : params)
            if (p.first == k) out.push_back(p.second);
        return out;
    }
};

struct ParseError {
    std::string msg;
};

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

std::string lower(std::string s)
{
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

std::string unquote(const std::string& s)
{
    if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
        return s.substr(1, s.size() - 2);
    return s;
}

[[noreturn]] void error_at(const Chunk& c, const std::string& what)
{
    throw ParseError{c.keyword + " (line " + std::to_string(c.line) + "): " + what};
}

bool parse_double(const std::string& s, double& v)
{
    const std::string t = trim(s);
    if (t.empty()) return false;
    char* end = nullptr;
    v = std::strtod(t.c_str(), &end);
    return end && *end == '\0';
}

double get_double(const Chunk& c, const std::string& k, double dflt)
{
    const std::string* v = c.get(k);
    if (!v) return dflt;
    double d;
    if (!parse_double(*v, d)) error_at(c, "`" + k + "` is not a number: " + *v);
    return d;
}

bool get_bool(const Chunk& c, const std::string& k, bool dflt)
{
    const std::string* v = c.get(k);
    if (!v) return dflt;
    const std::string t = lower(trim(*v));
    if (t == "true" || t == "1" || t == "yes") return true;
    if (t == "false" || t == "0" || t == "no") return false;
    error_at(c, "`" + k + "` is not a boolean: " + *v);
}

bool get_vec3(const Chunk& c, const std::string& k, double v[3])
{
    const std::string* s = c.get(k);
    if (!s) return false;
    std::istringstream is(*s);
    if (!(is >> v[0] >> v[1] >> v[2])) error_at(c, "`" + k + "` needs three numbers: " + *s);
    return true;
}

// Reads the chunks of a native-v7 scene file: a keyword on its own line,
// `{` and `}` on their own lines, one `key value...` parameter per line;
// `#` line comments and `/* */` block comments (Parsers/README.md).
std::vector<Chunk> read_chunks(const std::string& path)
{
    std::ifstream in(path);
    if (!in) throw ParseError{"cannot open " + path};
    std::vector<Chunk> chunks;
    std::string raw;
    int ln = 0;
    bool in_block = false, header = false, in_chunk = false, want_open = false;
    Chunk cur;
    while (std::getline(in, raw)) {
        ++ln;
        std::string s;
        for (size_t i = 0; i < raw.size(); ++i) {
            if (in_block) {
                if (raw.compare(i, 2, "*/") == 0) {
                    in_block = false;
                    ++i;
                }
                continue;
            }
            if (raw.compare(i, 2, "/*") == 0) {
                in_block = true;
                ++i;
                continue;
            }
            if (raw[i] == '#') break;
            s.push_back(raw[i]);
        }
        s = trim(s);
        if (s.empty()) continue;
        if (!header) {
            if (s.compare(0, 16, "RISE ASCII SCENE") != 0) throw ParseError{path + ": not a RISE ASCII SCENE file"};
            header = true;
            continue;
        }
        if (want_open) {
            if (s != "{") throw ParseError{"line " + std::to_string(ln) + ": expected `{` after " + cur.keyword};
            want_open = false;
            in_chunk = true;
            continue;
        }
        if (in_chunk) {
            if (s == "}") {
                chunks.push_back(cur);
                in_chunk = false;
                continue;
            }
            const size_t sp = s.find_first_of(" \t");
            const std::string key = sp == std::string::npos ? s : s.substr(0, sp);
            const std::string val = sp == std::string::npos ? std::string() : trim(s.substr(sp));
            if (val.compare(0, 5, "expr(") == 0)
                throw ParseError{"line " + std::to_string(ln) + ": `expr(...)` values are not supported"};
            cur.params.emplace_back(key, val);
            continue;
        }
        if (s[0] == '>') {
            // Job::LoadAsciiSceneViaCst accepts only the render-neutral
            // commands (`> echo`, `> set accelerator`, Job.h:2835-2842) and
            // refuses the rest; so does this reader.
            std::istringstream is(s.substr(1));
            std::string cmd, arg;
            is >> cmd >> arg;
            if (cmd == "echo" || (cmd == "set" && arg == "accelerator")) continue;
            throw ParseError{"line " + std::to_string(ln) + ": `> " + cmd +
                             "` commands are not supported (only `> echo` and `> set accelerator` are render-neutral)"};
        }
        if (s[0] == '!' || s[0] == '$')
            throw ParseError{"line " + std::to_string(ln) + ": macros of the retired scene language are not supported"};
        cur = Chunk{};
        cur.keyword = s;
        cur.line = ln;
        want_open = true;
    }
    if (in_chunk || want_open) throw ParseError{path + ": unterminated chunk " + cur.keyword};
    if (!header) throw ParseError{path + ": empty file"};
    return chunks;
}

std::string dir_of(const std::string& path)
{
    const size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string(".") : path.substr(0, k);
}

std::string resolve(const std::string& base_dir, const std::string& name)
{
    const std::string n = unquote(name);
    if (!n.empty() && n[0] == '/') return n;
    const std::string a = base_dir + "/" + n;
    if (std::ifstream(a).good()) return a;
    return n;  // relative to the working directory
}

// PiecewiseLinearFunction1D::getYValueFromFunction (Utilities/PiecewiseLinearFunction.h:55-90):
// clamped at both ends, linear in between, 1 with fewer than two points.
double piecewise_linear(const std::vector<std::pair<double, double>>& cp, double s)
{
    if (cp.size() < 2) return 1.0;
    if (s <= cp.front().first) return cp.front().second;
    if (s >= cp.back().first) return cp.back().second;
    auto it = std::lower_bound(cp.begin(), cp.end(), std::make_pair(s, -1e300),
                               [](const std::pair<double, double>& a, const std::pair<double, double>& b) {
                                   return a.first < b.first;
                               });
    if (it->first == s) return it->second;
    if ((it - 1)->first == s) return (it - 1)->second;
    const double d = (s - (it - 1)->first) / (it->first - (it - 1)->first);
    return (d * (it->second - (it - 1)->second)) + (it - 1)->second;
}

struct Builder {
    rise_hip_scene_file& F;
    std::string base_dir;
    std::map<std::string, int> spd_of_painter;
    struct RgbPainter {
        double rgb[3];  // linear Rec.709
        int spd = -1;   // uplifted spectrum, made on first use
    };
    std::map<std::string, RgbPainter> rgb_painters;
    std::map<double, int> flat_spd;
    std::map<std::string, int> material_of;
    std::map<std::string, int> medium_of;
    struct Mesh {
        std::vector<double> v;  // object space, 9 per triangle
        bool double_sided = false;
        RiseMeshPtr risemesh;   // the file of a risemesh_geometry (its BVH2 cache)
    };
    std::map<std::string, Mesh> meshes;
    bool have_camera = false, have_film = false, have_rasterizer = false;

    int add_spd(double begin, double end, std::vector<float> amp)
    {
        F.amps.push_back(std::move(amp));
        rise_hip_spd_desc d{};
        d.lambda_begin = (float)begin;
        d.lambda_end = (float)end;
        d.n = (int32_t)F.amps.back().size();
        d.amplitudes = nullptr;  // bound once all spectra exist
        F.spds.push_back(d);
        return (int)F.spds.size() - 1;
    }

    int flat(double v)
    {
        auto it = flat_spd.find(v);
        if (it != flat_spd.end()) return it->second;
        // A number stands for a uniform painter: the same value at every
        // wavelength, so one bin over a range wider than any render's.
        const int k = add_spd(0.0, 10000.0, std::vector<float>(1, (float)v));
        flat_spd[v] = k;
        return k;
    }

    // A material parameter: a spectral painter's name, or a number (a
    // uniform scalar painter).  "none" -> -1 when allowed.
    int spectrum_param(const Chunk& c, const std::string& key, const std::string& dflt, bool allow_none)
    {
        const std::string* pv = c.get(key);
        const std::string v = unquote(trim(pv ? *pv : dflt));
        if (v == "none" || v.empty()) {
            if (allow_none) return -1;
            error_at(c, "`" + key + "` is required");
        }
        double d;
        if (parse_double(v, d)) return flat(d);
        auto it = spd_of_painter.find(v);
        if (it != spd_of_painter.end()) return it->second;
        auto ri = rgb_painters.find(v);
        if (ri != rgb_painters.end()) {
            // RGBAlbedoSpectrum::FromRGB (UniformColorPainter.h:38-52): the
            // sigmoid spectrum of the table entry, scale 1.
            if (ri->second.spd < 0) {
                const rise_hip_rgb_lut* lut = nullptr;
                float cf[3];
                if (rise_hip_rgb_lut_default(&lut) != RISE_HIP_OK ||
                    rise_hip_rgb_lut_lookup(lut, ri->second.rgb, cf) != RISE_HIP_OK)
                    error_at(c, std::string(rise_hip_last_error()));
                F.amps.push_back({cf[0], cf[1], cf[2], 1.0f});
                rise_hip_spd_desc d{};
                d.lambda_begin = 380.0f;
                d.lambda_end = 780.0f;
                d.n = RISE_HIP_SPD_SIGMOID;
                F.spds.push_back(d);
                ri->second.spd = (int)F.spds.size() - 1;
            }
            return ri->second.spd;
        }
        error_at(c, "`" + key + "` names an unknown painter `" + v + "`");
    }

    // :1005-1040 (uniformcolor_painter) and Job::AddUniformColorPainter
    // (Job.cpp:2454-2485): `color` R G B in `colorspace` Rec709RGB_Linear
    // (default; RISEPel is linear Rec.709), RISERGB (the same) or sRGB
    // (decoded by ColorUtils::SRGBTransferFunctionInverse, ColorUtils.cpp:34-45).
    void uniform_painter(const Chunk& c)
    {
        double col[3] = {0, 0, 0};
        get_vec3(c, "color", col);
        const std::string cs = unquote(c.get("colorspace") ? *c.get("colorspace") : "Rec709RGB_Linear");
        RgbPainter p;
        for (int k = 0; k < 3; ++k) {
            double v = col[k];
            if (cs == "sRGB") v = v == 0 ? 0.0 : v <= 0.03928 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4);
            else if (cs != "Rec709RGB_Linear" && cs != "RISERGB")
                error_at(c, "colorspace `" + cs + "` is not supported (Rec709RGB_Linear, RISERGB or sRGB)");
            p.rgb[k] = v;
        }
        rgb_painters[unquote(c.get("name") ? *c.get("name") : "noname")] = p;
    }

    // Parsers/ChunkParserRegistry.cpp:1086-1170 and Job::AddSpectralColorPainter
    // (Job.cpp:1557-1582): a SpectralPacket of n = number of control points
    // bins over [nmbegin, nmend], bin i holding the piecewise-linear
    // interpolation of the points at nmbegin + i * (nmend - nmbegin) / n.
    // The painter's `scale` only scales its RGB colour
    // (Painters/SpectralColorPainter.cpp:28-44), not the spectrum the
    // spectral renderer reads.
    void spectral_painter(const Chunk& c)
    {
        const std::string name = unquote(c.get("name") ? *c.get("name") : "noname");
        const double nmbegin = get_double(c, "nmbegin", 400.0), nmend = get_double(c, "nmend", 700.0);
        std::vector<double> nms, amps;
        for (const std::string& cp : c.all("cp")) {
            double a = 0, b = 0;
            if (std::sscanf(cp.c_str(), "%lf %lf", &a, &b) < 1) error_at(c, "bad `cp` " + cp);
            nms.push_back(a);
            amps.push_back(b);
        }
        auto read_list = [&](const std::string& key, int cols, std::vector<double>& x, std::vector<double>* y) {
            if (!c.has(key)) return;
            const std::string fn = resolve(base_dir, *c.get(key));
            std::ifstream f(fn);
            if (!f) error_at(c, "cannot open `" + fn + "`");
            std::string line;
            while (std::getline(f, line)) {
                const std::string t = trim(line);
                if (t.empty() || t[0] == '#') continue;
                double a = 0, b = 0;
                const int got = std::sscanf(t.c_str(), "%lf %lf", &a, &b);
                if (cols == 2 && got == 2) {
                    x.push_back(a);
                    y->push_back(b);
                } else if (cols == 1 && got >= 1) {
                    x.push_back(a);
                }
            }
        };
        read_list("file", 2, nms, &amps);
        read_list("nmfile", 1, nms, nullptr);
        read_list("ampfile", 1, amps, nullptr);
        if (nms.empty() || amps.empty()) error_at(c, "spectral painter `" + name + "` has no samples");
        const size_t n = std::min(nms.size(), amps.size());
        std::vector<std::pair<double, double>> pts;
        for (size_t i = 0; i < n; ++i) pts.emplace_back(nms[i], amps[i]);
        std::stable_sort(pts.begin(), pts.end(),
                         [](const std::pair<double, double>& a, const std::pair<double, double>& b) {
                             return a.first < b.first;
                         });
        if (!(nmend > nmbegin)) error_at(c, "nmend must exceed nmbegin");
        const double delta = (nmend - nmbegin) / (double)n;
        std::vector<float> bins(n);
        for (size_t i = 0; i < n; ++i) bins[i] = (float)piecewise_linear(pts, nmbegin + (double)i * delta);
        spd_of_painter[name] = add_spd(nmbegin, nmend, std::move(bins));
    }

    // Parsers/ChunkParserRegistry.cpp:3090 (LambertianMaterial).
    void lambertian(const Chunk& c)
    {
        rise_hip_material_desc m{};
        m.type = RISE_HIP_MATERIAL_LAMBERTIAN;
        m.reflectance_spd = spectrum_param(c, "reflectance", "none", false);
        m.emission_spd = -1;
        m.emission_scale = 1.0f;
        m.specular_spd = m.ior_spd = m.extinction_spd = -1;
        add_material(c, m);
    }

    // :3373 (LambertianLuminaireMaterial): exitance painter times scale over
    // the reflective `material` (none: black).
    void luminaire(const Chunk& c)
    {
        rise_hip_material_desc m{};
        m.type = RISE_HIP_MATERIAL_LAMBERTIAN;
        const std::string base = unquote(c.get("material") ? *c.get("material") : "none");
        if (base == "none") {
            m.reflectance_spd = flat(0.0);
        } else {
            auto it = material_of.find(base);
            if (it == material_of.end()) error_at(c, "unknown material `" + base + "`");
            const rise_hip_material_desc& b = F.materials[(size_t)it->second];
            if (b.type != RISE_HIP_MATERIAL_LAMBERTIAN || b.emission_spd >= 0)
                error_at(c, "a luminaire's reflective material must be a plain lambertian_material");
            m.reflectance_spd = b.reflectance_spd;
        }
        m.emission_spd = spectrum_param(c, "exitance", "none", false);
        m.emission_scale = (float)get_double(c, "scale", 1.0);
        m.specular_spd = m.ior_spd = m.extinction_spd = -1;
        add_material(c, m);
    }

    // :3826 (GGXMaterial): conductor Fresnel, isotropic roughness, no emission.
    void ggx(const Chunk& c)
    {
        const std::string mode = lower(unquote(c.get("fresnel_mode") ? *c.get("fresnel_mode") : "conductor"));
        if (mode != "conductor") error_at(c, "only fresnel_mode conductor is supported");
        for (const char* k : {"emissive", "film_ior", "film_extinction", "film_thickness", "tangent_rotation"})
            if (c.has(k) && unquote(*c.get(k)) != "none") error_at(c, std::string("`") + k + "` is not supported");
        double ax = 0.15, ay = 0.15;
        if (c.has("alphax") && !parse_double(*c.get("alphax"), ax)) error_at(c, "`alphax` must be a number");
        if (c.has("alphay") && !parse_double(*c.get("alphay"), ay)) error_at(c, "`alphay` must be a number");
        if (ax != ay) error_at(c, "anisotropic roughness (alphax != alphay) is not supported");
        rise_hip_material_desc m{};
        m.type = RISE_HIP_MATERIAL_GGX;
        m.reflectance_spd = spectrum_param(c, "rd", "none", true);
        m.emission_spd = -1;
        m.emission_scale = 1.0f;
        m.specular_spd = spectrum_param(c, "rs", "none", false);
        m.ior_spd = spectrum_param(c, "ior", "2.45", false);
        m.extinction_spd = spectrum_param(c, "extinction", "3.45", false);
        m.alpha = (float)ax;
        add_material(c, m);
    }

    // :3200 (DielectricMaterial), the smooth case only.
    void dielectric(const Chunk& c)
    {
        const double scat = get_double(c, "scattering", 10000.0);
        if (scat < 1e6) error_at(c, "only smooth dielectrics (scattering >= 1e6) are supported");
        if (get_bool(c, "henyey-greenstein", false) || c.has("ar_layer"))
            error_at(c, "Henyey-Greenstein scattering and coating layers are not supported");
        rise_hip_material_desc m{};
        m.type = RISE_HIP_MATERIAL_DIELECTRIC;
        m.reflectance_spd = -1;
        m.emission_spd = -1;
        m.emission_scale = 1.0f;
        m.specular_spd = -1;
        m.ior_spd = spectrum_param(c, "ior", "1.33", false);
        m.extinction_spd = spectrum_param(c, "tau", "none", true);
        add_material(c, m);
    }

    void add_material(const Chunk& c, const rise_hip_material_desc& m)
    {
        const std::string name = unquote(c.get("name") ? *c.get("name") : "noname");
        F.materials.push_back(m);
        material_of[name] = (int)F.materials.size() - 1;
    }

    // :6142 (HomogeneousMedium): RGB coefficients, phase "hg <g>" or "isotropic".
    void medium(const Chunk& c)
    {
        if (c.has("absorption_spectral") || c.has("scattering_spectral") || c.has("emission"))
            error_at(c, "spectral coefficients and emission are not supported");
        double sa[3] = {0, 0, 0}, ss[3] = {0, 0, 0};
        get_vec3(c, "absorption", sa);
        get_vec3(c, "scattering", ss);
        double g = 0.0;
        if (c.has("phase")) {
            std::istringstream is(lower(unquote(*c.get("phase"))));
            std::string kind;
            is >> kind;
            if (kind == "hg") {
                if (!(is >> g)) error_at(c, "`phase hg` needs g");
            } else if (kind != "isotropic") {
                error_at(c, "phase `" + kind + "` is not supported (hg or isotropic)");
            }
        }
        rise_hip_medium_desc m{};
        for (int k = 0; k < 3; ++k) {
            m.sigma_a[k] = (float)sa[k];
            m.sigma_s[k] = (float)ss[k];
        }
        m.hg_g = (float)g;
        F.media.push_back(m);
        medium_of[unquote(c.get("name") ? *c.get("name") : "noname")] = (int)F.media.size() - 1;
    }

    // :6931 (Job::AddHosekWilkieSkylight): the radiance map and, unless
    // create_sun is FALSE, its matched directional sun light.  ground_albedo has
    // no effect on the model (HosekWilkieSkyModel.cpp:135-155) and is ignored.
    void sky(const Chunk& c)
    {
        F.has_sky = true;
        F.sky.solar_elevation_deg = (float)get_double(c, "solar_elevation", 45.0);
        F.sky.solar_azimuth_deg = (float)get_double(c, "solar_azimuth", 0.0);
        F.sky.turbidity = (float)get_double(c, "turbidity", 3.0);
        F.sky.sky_scale = (float)get_double(c, "sky_intensity_scale", 1.0);
        F.sky.create_sun = get_bool(c, "create_sun", true) ? 1 : 0;
        F.sky.sun_intensity_scale = (float)get_double(c, "sun_intensity_scale", 3.14);
    }

    // :6976 (Job::AddDirectionalLight): power (1), direction and color (0 0 0).
    void directional_light(const Chunk& c)
    {
        rise_hip_directional_light_desc d{};
        double dir[3] = {0, 0, 0}, col[3] = {0, 0, 0};
        get_vec3(c, "direction", dir);
        get_vec3(c, "color", col);
        if (!(dir[0] != 0 || dir[1] != 0 || dir[2] != 0)) error_at(c, "`direction` is zero");
        for (int k = 0; k < 3; ++k) {
            d.direction[k] = (float)dir[k];
            d.color[k] = (float)col[k];
        }
        d.power = (float)get_double(c, "power", 1.0);
        F.dir_lights.push_back(d);
    }

    // :5104 and Geometry/TriangleMeshLoaderRAW2.cpp: "nverts ntris", then
    // "v x y z nx ny nz u v [r g b]" lines, then "t a b c" lines.  The GPU
    // path shades with face normals.
    void rawmesh2(const Chunk& c)
    {
        const std::string name = unquote(c.get("name") ? *c.get("name") : "noname");
        const std::string fn = resolve(base_dir, c.get("file") ? *c.get("file") : "none");
        std::ifstream f(fn);
        if (!f) error_at(c, "cannot open `" + fn + "`");
        std::string line;
        unsigned nv = 0, nt = 0;
        if (!std::getline(f, line) || std::sscanf(line.c_str(), "%u %u", &nv, &nt) != 2)
            error_at(c, "`" + fn + "`: bad header");
        std::vector<double> v((size_t)nv * 3);
        for (unsigned i = 0; i < nv; ++i) {
            char type = 0;
            double x, y, z;
            if (!std::getline(f, line) || std::sscanf(line.c_str(), " %c %lf %lf %lf", &type, &x, &y, &z) != 4 ||
                type != 'v')
                error_at(c, "`" + fn + "`: expected vertex " + std::to_string(i));
            v[3 * i] = x;
            v[3 * i + 1] = y;
            v[3 * i + 2] = z;
        }
        Mesh m;
        m.double_sided = get_bool(c, "double_sided", false);
        m.v.reserve((size_t)nt * 9);
        for (unsigned i = 0; i < nt; ++i) {
            char type = 0;
            unsigned a, b, d;
            if (!std::getline(f, line) || std::sscanf(line.c_str(), " %c %u %u %u", &type, &a, &b, &d) != 4 ||
                type != 't' || a >= nv || b >= nv || d >= nv)
                error_at(c, "`" + fn + "`: bad triangle " + std::to_string(i));
            for (unsigned k : {a, b, d})
                for (int j = 0; j < 3; ++j) m.v.push_back(v[3 * k + j]);
        }
        meshes[name] = std::move(m);
    }

    // :5137 (risemesh_geometry, Job::AddRISEMeshTriangleMeshGeometry): a
    // .risemesh file; double-sidedness comes from the file, `face_normals`
    // is what the GPU path shades with anyway.  The file's BVH2 cache, when
    // valid, replaces the SAH build of this object's subtree.
    void risemesh(const Chunk& c)
    {
        const std::string name = unquote(c.get("name") ? *c.get("name") : "noname");
        const std::string fn = resolve(base_dir, c.get("file") ? *c.get("file") : "none");
        rise_hip_risemesh* raw = nullptr;
        if (rise_hip_risemesh_read(fn.c_str(), &raw) != RISE_HIP_OK)
            error_at(c, std::string(rise_hip_last_error()));
        Mesh m;
        m.risemesh = RiseMeshPtr(raw, RiseMeshDeleter());
        m.v = risemesh_triangles_f64(raw);
        m.double_sided = risemesh_double_sided(raw);
        meshes[name] = std::move(m);
    }

    // :5004 (ClippedPlaneGeometry), stored as the triangles (a, b, c) and
    // (a, c, d); its `doublesided` defaults to TRUE.
    void clippedplane(const Chunk& c)
    {
        if (c.has("file")) error_at(c, "file-based clipped planes are not supported");
        double p[4][3] = {};
        const char* keys[4] = {"pta", "ptb", "ptc", "ptd"};
        for (int k = 0; k < 4; ++k)
            if (!get_vec3(c, keys[k], p[k])) error_at(c, std::string("`") + keys[k] + "` is required");
        Mesh m;
        m.double_sided = get_bool(c, "doublesided", true);
        for (int t : {0, 1, 2, 0, 2, 3})
            for (int j = 0; j < 3; ++j) m.v.push_back(p[t][j]);
        meshes[unquote(c.get("name") ? *c.get("name") : "noname")] = std::move(m);
    }

    // :6411 (Job::AddObject, Job.cpp:5576-5640; Transformable::Finalize-
    // Transformations): world = Position * Orientation * Stretch in RISE's
    // row-vector matrices, i.e. scale, then rotate about z, y, x
    // (orientation in degrees), then translate.
    void object(const Chunk& c)
    {
        const std::string name = unquote(c.get("name") ? *c.get("name") : "noname");
        if (c.has("matrix") || c.has("quaternion")) error_at(c, "`matrix` / `quaternion` transforms are not supported");
        if (!get_bool(c, "receives_shadows", true)) error_at(c, "receives_shadows FALSE is not supported");
        for (const char* k : {"modifier", "shader", "radiance_map"})
            if (c.has(k) && unquote(*c.get(k)) != "none") error_at(c, std::string("`") + k + "` is not supported");
        const std::string geo = unquote(c.get("geometry") ? *c.get("geometry") : "none");
        auto gi = meshes.find(geo);
        if (gi == meshes.end()) error_at(c, "unknown or unsupported geometry `" + geo + "`");
        const std::string mat = unquote(c.get("material") ? *c.get("material") : "none");
        auto mi = material_of.find(mat);
        if (mi == material_of.end()) error_at(c, "unknown material `" + mat + "`");
        double pos[3] = {0, 0, 0}, ori[3] = {0, 0, 0}, scl[3] = {1, 1, 1};
        get_vec3(c, "position", pos);
        get_vec3(c, "orientation", ori);
        get_vec3(c, "scale", scl);
        for (double& a : ori) a *= kDegToRad;
        const double cx = std::cos(ori[0]), sx = std::sin(ori[0]), cy = std::cos(ori[1]), sy = std::sin(ori[1]),
                     cz = std::cos(ori[2]), sz = std::sin(ori[2]);
        const std::vector<double>& src = gi->second.v;
        std::vector<float> w(src.size());
        for (size_t i = 0; i < src.size(); i += 3) {
            double x = src[i] * scl[0], y = src[i + 1] * scl[1], z = src[i + 2] * scl[2];
            double t;
            t = cz * x - sz * y;  // ZRotation (MatricesOps.h:305-319), row vector
            y = sz * x + cz * y;
            x = t;
            t = cy * x + sy * z;  // YRotation (:290-303)
            z = -sy * x + cy * z;
            x = t;
            t = cx * y - sx * z;  // XRotation (:274-287)
            z = sx * y + cx * z;
            y = t;
            w[i] = (float)(x + pos[0]);
            w[i + 1] = (float)(y + pos[1]);
            w[i + 2] = (float)(z + pos[2]);
        }
        F.meshes.push_back(std::move(w));
        F.object_risemesh.push_back(gi->second.risemesh);
        rise_hip_object_desc od{};
        od.n_tris = (int64_t)(F.meshes.back().size() / 9);
        od.material = mi->second;
        od.double_sided = gi->second.double_sided ? 1 : 0;
        F.objects.push_back(od);
        rise_hip_object_media om{};
        om.interior_medium = -1;
        const std::string med = unquote(c.get("interior_medium") ? *c.get("interior_medium") : "none");
        if (med != "none") {
            auto it = medium_of.find(med);
            if (it == medium_of.end()) error_at(c, "unknown medium `" + med + "`");
            om.interior_medium = it->second;
            F.any_media = true;
        }
        om.casts_shadows = get_bool(c, "casts_shadows", true) ? 1 : 0;
        if (!om.casts_shadows) F.any_media = true;
        F.object_media.push_back(om);
        (void)name;
    }

    // :4282 (PinholeCamera): location, lookat, up, fov (degrees, default 30).
    void camera(const Chunk& c)
    {
        for (const char* k : {"orientation", "target_orientation", "theta", "phi", "pitch", "yaw", "roll", "fstop"})
            if (c.has(k) && get_double(c, k, 0.0) != 0.0) error_at(c, std::string("`") + k + "` is not supported");
        double loc[3] = {0, 0, 0}, at[3] = {0, 0, -1}, up[3] = {0, 1, 0};
        get_vec3(c, "location", loc);
        get_vec3(c, "lookat", at);
        get_vec3(c, "up", up);
        for (int k = 0; k < 3; ++k) {
            F.desc.camera.position[k] = (float)loc[k];
            F.desc.camera.look_at[k] = (float)at[k];
            F.desc.camera.up[k] = (float)up[k];
        }
        F.desc.camera.fov_deg = (float)get_double(c, "fov", 30.0);
        have_camera = true;
    }

    // :4814 (film): width, height (default 960 x 540, Rendering/Film.h:37-38).
    void film(const Chunk& c)
    {
        F.desc.width = (int32_t)get_double(c, "width", 960.0);
        F.desc.height = (int32_t)get_double(c, "height", 540.0);
        have_film = true;
    }

    // :8894 (PathTracingSpectralRasterizer): samples (default 32), nmbegin /
    // nmend (380 / 780), Russian roulette (3, 0.05), bounce caps (unlimited:
    // the 128-vertex path limit bounds them), radiance_background.  The GPU
    // renders with hero-wavelength sampling whether or not `hwss` is set
    // (the same expectation as single-wavelength sampling); the pixel filter
    // is the box filter; denoising and adaptive sampling are not applied.
    void rasterizer(const Chunk& c)
    {
        const int spp = (int)get_double(c, "samples", 32.0);
        F.desc.spp = F.desc.spp_total = spp;
        F.desc.spp_offset = 0;
        F.desc.lambda_begin = (float)get_double(c, "nmbegin", 380.0);
        F.desc.lambda_end = (float)get_double(c, "nmend", 780.0);
        F.desc.rr_min_depth = (int32_t)get_double(c, "rr_min_depth", 3.0);
        F.desc.rr_threshold = (float)get_double(c, "rr_threshold", 0.05);
        auto cap = [&](const char* k) {
            const double v = get_double(c, k, 255.0);
            return (int32_t)std::min(255.0, std::max(0.0, v));
        };
        F.desc.max_diffuse_bounce = cap("max_diffuse_bounce");
        F.desc.max_glossy_bounce = cap("max_glossy_bounce");
        F.desc.max_transmission_bounce = cap("max_transmission_bounce");
        F.desc.sky_background = get_bool(c, "radiance_background", true) ? 1 : 0;
        // PixelFilterConfig (Utilities/PixelFilterConfig.h:80-97): the Sobol pixel
        // sampler, with the Z-order (blue-noise) index unless blue_noise_sampler is off.
        F.desc.sampler = get_bool(c, "blue_noise_sampler", true) ? RISE_HIP_SAMPLER_ZSOBOL : RISE_HIP_SAMPLER_SOBOL;
        if (c.has("pixel_sampler") && lower(unquote(*c.get("pixel_sampler"))) != "sobol")
            error_at(c, "only the sobol pixel_sampler is supported");
        const std::string pf = lower(unquote(c.get("pixel_filter") ? *c.get("pixel_filter") : "box"));
        if (pf != "box") error_at(c, "only the box pixel filter is supported");
        if (c.has("radiance_map")) error_at(c, "`radiance_map` is not supported (use hosek_wilkie_skylight)");
        have_rasterizer = true;
    }

    // :9207-9325 (FileRasterizerOutput): pattern, type (TGA default), bpp (8),
    // color_space (sRGB), exposure, display_transform (ACES for LDR types,
    // none for HDR), exr_with_alpha (TRUE).  `multiple` (one file per
    // animation frame) has no meaning for a single render and is ignored.
    void file_output(const Chunk& c)
    {
        rise_hip_output_desc d{};
        const std::string t = unquote(c.get("type") ? *c.get("type") : "TGA");
        if (t == "TGA") d.format = RISE_HIP_IMAGE_TGA;
        else if (t == "PPM") d.format = RISE_HIP_IMAGE_PPM;
        else if (t == "PNG") d.format = RISE_HIP_IMAGE_PNG;
        else if (t == "HDR") d.format = RISE_HIP_IMAGE_HDR;
        else if (t == "EXR") d.format = RISE_HIP_IMAGE_EXR;
        else error_at(c, "output type `" + t + "` is not supported (TGA, PPM, PNG, HDR, EXR)");
        d.bpp = (int32_t)get_double(c, "bpp", 8.0);
        const std::string cs = unquote(c.get("color_space") ? *c.get("color_space") : "sRGB");
        if (cs == "sRGB") d.color_space = RISE_HIP_COLOR_SRGB;
        else if (cs == "Rec709RGB_Linear") d.color_space = RISE_HIP_COLOR_REC709_LINEAR;
        else error_at(c, "color_space `" + cs + "` is not supported (sRGB, Rec709RGB_Linear)");
        d.exposure_ev = get_double(c, "exposure", 0.0);
        d.display_transform = RISE_HIP_DISPLAY_DEFAULT;
        if (c.has("display_transform")) {
            const std::string dt = unquote(*c.get("display_transform"));
            static const char* names[] = {"none", "reinhard", "aces", "agx", "hable"};
            d.display_transform = -2;
            for (int k = 0; k < 5; ++k)
                if (dt == names[k]) d.display_transform = k;
            if (d.display_transform == -2) error_at(c, "unknown display_transform `" + dt + "`");
        }
        d.with_alpha = get_bool(c, "exr_with_alpha", true) ? 1 : 0;
        F.outputs.push_back(d);
        F.output_patterns.push_back(unquote(c.get("pattern") ? *c.get("pattern") : "none"));
    }

    void run(const std::vector<Chunk>& chunks)
    {
        // Defaults of a render the file does not fully describe.
        F.desc.width = 960;
        F.desc.height = 540;
        F.desc.spp = F.desc.spp_total = 32;
        F.desc.lambda_begin = 380.0f;
        F.desc.lambda_end = 780.0f;
        F.desc.max_diffuse_bounce = F.desc.max_glossy_bounce = F.desc.max_transmission_bounce = 255;
        F.desc.rr_min_depth = 3;
        F.desc.rr_threshold = 0.05f;
        F.desc.seed = 1;
        F.desc.camera.look_at[2] = -1.0f;
        F.desc.camera.up[1] = 1.0f;
        F.desc.camera.fov_deg = 30.0f;
        F.desc.sky_background = 1;
        static const char* kSkip[] = {"standard_shader", "pathtracing_shaderop",
                                      "directlighting_shaderop", "ambientocclusion_shaderop", "simple_shader",
                                      "advanced_shader", "global_options", "pixelpel_rasterizer"};
        for (const Chunk& c : chunks) {
            const std::string& k = c.keyword;
            if (k == "spectral_painter") spectral_painter(c);
            else if (k == "uniformcolor_painter") uniform_painter(c);
            else if (k == "lambertian_material") lambertian(c);
            else if (k == "lambertian_luminaire_material") luminaire(c);
            else if (k == "ggx_material") ggx(c);
            else if (k == "dielectric_material") dielectric(c);
            else if (k == "homogeneous_medium") medium(c);
            else if (k == "hosek_wilkie_skylight") sky(c);
            else if (k == "directional_light") directional_light(c);
            else if (k == "rawmesh2_geometry") rawmesh2(c);
            else if (k == "clippedplane_geometry") clippedplane(c);
            else if (k == "risemesh_geometry") risemesh(c);
            else if (k == "standard_object") object(c);
            else if (k == "pinhole_camera") camera(c);
            else if (k == "film") film(c);
            else if (k == "pathtracing_spectral_rasterizer") rasterizer(c);
            else if (k == "file_rasterizeroutput") file_output(c);
            else if (std::find_if(std::begin(kSkip), std::end(kSkip), [&](const char* s) { return k == s; }) !=
                         std::end(kSkip) ||
                     k.find("shaderop") != std::string::npos)
                ++F.skipped;
            else
                error_at(c, "chunk not supported by the GPU path tracer");
        }
        if (F.objects.empty()) throw ParseError{"the scene has no objects"};
        if (!have_camera) throw ParseError{"the scene has no pinhole_camera"};
        for (size_t i = 0; i < F.spds.size(); ++i) F.spds[i].amplitudes = F.amps[i].data();
        for (size_t i = 0; i < F.objects.size(); ++i) F.objects[i].vertices = F.meshes[i].data();
    }
};

}  // namespace

extern "C" rise_hip_status rise_hip_scene_file_parse(const char* path, rise_hip_scene_file** out_file)
{
    clear_error();
    if (!path || !out_file) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out_file = nullptr;
    rise_hip_scene_file* f = new (std::nothrow) rise_hip_scene_file;
    if (!f) return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    try {
        const std::vector<Chunk> chunks = read_chunks(path);
        Builder b{*f, dir_of(path)};
        b.run(chunks);
    } catch (const ParseError& e) {
        delete f;
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string("scene file: ") + e.msg);
    } catch (const std::bad_alloc&) {
        delete f;
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed while reading the scene file");
    }
    *out_file = f;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_scene_file_get_info(const rise_hip_scene_file* f, rise_hip_scene_file_info* out)
{
    clear_error();
    if (!f || !out) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = rise_hip_scene_file_info{};
    out->n_objects = (int64_t)f->objects.size();
    out->n_materials = (int64_t)f->materials.size();
    out->n_spds = (int64_t)f->spds.size();
    out->n_media = (int64_t)f->media.size();
    for (const auto& o : f->objects) out->n_tris += o.n_tris;
    out->has_sky = f->has_sky ? 1 : 0;
    out->n_skipped_chunks = f->skipped;
    out->n_outputs = (int32_t)f->outputs.size();
    out->n_directional_lights = (int32_t)f->dir_lights.size();
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_scene_file_get_directional_lights(const rise_hip_scene_file* f,
                                                                      const rise_hip_directional_light_desc** lights)
{
    clear_error();
    if (!f || !lights) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *lights = f->dir_lights.data();
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_scene_file_get_tables(const rise_hip_scene_file* f,
                                                          const rise_hip_object_desc** objects,
                                                          const rise_hip_object_media** object_media,
                                                          const rise_hip_material_desc** materials,
                                                          const rise_hip_spd_desc** spds,
                                                          const rise_hip_medium_desc** media,
                                                          const rise_hip_sky_desc** sky)
{
    clear_error();
    if (!f) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL scene file");
    if (objects) *objects = f->objects.data();
    if (object_media) *object_media = f->object_media.data();
    if (materials) *materials = f->materials.data();
    if (spds) *spds = f->spds.data();
    if (media) *media = f->media.data();
    if (sky) *sky = f->has_sky ? &f->sky : nullptr;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_scene_file_get_render_desc(const rise_hip_scene_file* f,
                                                               rise_hip_render_desc* out)
{
    clear_error();
    if (!f || !out) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = f->desc;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_scene_file_get_output(const rise_hip_scene_file* f, int32_t index,
                                                          rise_hip_output_desc* desc, const char** pattern)
{
    clear_error();
    if (!f) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL scene file");
    if (index < 0 || (size_t)index >= f->outputs.size())
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "output index out of range");
    if (desc) *desc = f->outputs[(size_t)index];
    if (pattern) *pattern = f->output_patterns[(size_t)index].c_str();
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_scene_file_upload(const rise_hip_scene_file* f, rise_hip_pt_scene** out_scene)
{
    clear_error();
    if (!f || !out_scene) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    std::vector<const Bvh2Tree*> trees(f->objects.size(), nullptr);
    for (size_t o = 0; o < trees.size(); ++o)
        if (f->object_risemesh[o]) trees[o] = risemesh_tree(f->object_risemesh[o].get());
    rise_hip_status st = pt_scene_create(f->objects.data(), f->objects.size(), f->materials.data(),
                                         f->materials.size(), f->spds.data(), f->spds.size(), f->media.data(),
                                         f->media.size(), f->any_media ? f->object_media.data() : nullptr,
                                         trees.data(), out_scene);
    if (st != RISE_HIP_OK) return st;
    if (!f->dir_lights.empty())
        st = rise_hip_pt_scene_set_directional_lights(*out_scene, f->dir_lights.data(), f->dir_lights.size());
    if (st == RISE_HIP_OK && f->has_sky) st = rise_hip_pt_scene_set_sky(*out_scene, &f->sky);
    if (st != RISE_HIP_OK) {
        rise_hip_pt_scene_release(*out_scene);
        *out_scene = nullptr;
    }
    return st;
}

extern "C" rise_hip_status rise_hip_scene_file_release(rise_hip_scene_file* f)
{
    clear_error();
    delete f;
    return RISE_HIP_OK;
}
