// json_scene.cpp -- the JSON scene reader of the reference-named C++ API.
//
// The parser is a small recursive-descent reader of RFC 8259 JSON (objects,
// arrays, strings with escapes, numbers, true/false/null).  The mapping from
// the document to FlipSolverParameters and scene objects follows
// Utils/jsonscenereader.cpp:79-136 and 183-300 field by field, defaults
// included.
#include "json_scene.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

namespace fghip {

namespace {

[[noreturn]] void bad(const std::string& msg) { throw Error(FG_E_INVALID, msg); }

class Parser {
public:
    explicit Parser(const std::string& t) : t_(t) {}

    JsonValue document()
    {
        JsonValue v = value();
        ws();
        if (p_ != t_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& t_;
    size_t p_ = 0;

    [[noreturn]] void fail(const char* what)
    {
        bad(std::string("JSON: ") + what + " at offset " + std::to_string(p_));
    }
    void ws()
    {
        while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\n' || t_[p_] == '\r')) p_++;
    }
    char peek()
    {
        ws();
        if (p_ >= t_.size()) fail("unexpected end");
        return t_[p_];
    }
    void expect(char c)
    {
        if (peek() != c) fail("unexpected character");
        p_++;
    }
    bool word(const char* w)
    {
        const size_t n = std::char_traits<char>::length(w);
        if (t_.compare(p_, n, w) != 0) return false;
        p_ += n;
        return true;
    }

    JsonValue value()
    {
        const char c = peek();
        JsonValue v;
        if (c == '{') {
            v.kind = JsonValue::OBJECT;
            p_++;
            if (peek() == '}') {
                p_++;
                return v;
            }
            for (;;) {
                if (peek() != '"') fail("expected a key");
                std::string key = string();
                expect(':');
                v.obj[key] = value();
                const char d = peek();
                p_++;
                if (d == '}') return v;
                if (d != ',') fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = JsonValue::ARRAY;
            p_++;
            if (peek() == ']') {
                p_++;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                const char d = peek();
                p_++;
                if (d == ']') return v;
                if (d != ',') fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JsonValue::STRING;
            v.str = string();
            return v;
        }
        if (word("true")) {
            v.kind = JsonValue::BOOL;
            v.b = true;
            return v;
        }
        if (word("false")) {
            v.kind = JsonValue::BOOL;
            return v;
        }
        if (word("null")) return v;
        return number();
    }

    std::string string()
    {
        expect('"');
        std::string out;
        while (p_ < t_.size() && t_[p_] != '"') {
            char c = t_[p_++];
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p_ >= t_.size()) fail("bad escape");
            c = t_[p_++];
            switch (c) {
            case '"': case '\\': case '/': out += c; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                if (p_ + 4 > t_.size()) fail("bad \\u escape");
                const unsigned cp = (unsigned)std::strtoul(t_.substr(p_, 4).c_str(), nullptr, 16);
                p_ += 4;
                // UTF-8 of a BMP code point (scene files are ASCII in practice)
                if (cp < 0x80) {
                    out += (char)cp;
                } else if (cp < 0x800) {
                    out += (char)(0xC0 | (cp >> 6));
                    out += (char)(0x80 | (cp & 0x3F));
                } else {
                    out += (char)(0xE0 | (cp >> 12));
                    out += (char)(0x80 | ((cp >> 6) & 0x3F));
                    out += (char)(0x80 | (cp & 0x3F));
                }
                break;
            }
            default: fail("bad escape");
            }
        }
        if (p_ >= t_.size()) fail("unterminated string");
        p_++;
        return out;
    }

    JsonValue number()
    {
        const size_t start = p_;
        if (p_ < t_.size() && (t_[p_] == '-' || t_[p_] == '+')) p_++;
        while (p_ < t_.size() && (isdigit((unsigned char)t_[p_]) || t_[p_] == '.' || t_[p_] == 'e' || t_[p_] == 'E' ||
                                  t_[p_] == '-' || t_[p_] == '+'))
            p_++;
        if (p_ == start) fail("unexpected character");
        const std::string s = t_.substr(start, p_ - start);
        char* end = nullptr;
        const double d = std::strtod(s.c_str(), &end);
        if (end != s.c_str() + s.size()) fail("bad number");
        JsonValue v;
        v.kind = JsonValue::NUMBER;
        v.num = d;
        return v;
    }
};

// tryGetValue (jsonscenereader.cpp:296-300): the value, or the default
double num_or(const JsonValue& o, const char* key, double dflt) { return o.has(key) ? o.at(key).number() : dflt; }
bool bool_or(const JsonValue& o, const char* key, bool dflt) { return o.has(key) ? o.at(key).boolean() : dflt; }
std::pair<float, float> pair_or(const JsonValue& o, const char* key, std::pair<float, float> dflt)
{
    if (!o.has(key)) return dflt;
    const JsonValue& a = o.at(key);
    if (a.kind != JsonValue::ARRAY || a.arr.size() != 2) bad(std::string("'") + key + "' must be a pair");
    return {(float)a.arr[0].number(), (float)a.arr[1].number()};
}

Geometry2d polygon(const JsonValue& o, float scale)
{
    const JsonValue& vs = o.at("verts");
    if (vs.kind != JsonValue::ARRAY) bad("'verts' must be an array");
    Geometry2d g;
    for (const JsonValue& v : vs.arr) {
        if (v.kind != JsonValue::ARRAY || v.arr.size() != 2) bad("a vertex must be a pair");
        g.addVertex(scale * Vec3((float)v.arr[0].number(), (float)v.arr[1].number()));
    }
    return g;
}

// emitterFromJson (jsonscenereader.cpp:197-226)
Emitter emitter(const JsonValue& o, float scale)
{
    Geometry2d g = polygon(o, scale);
    Emitter e(g);
    e.setTemperature((float)num_or(o, "temperature", 273.f));
    e.setConcentrartion((float)num_or(o, "concentration", 1.f));
    e.setViscosity((float)num_or(o, "viscosity", 0.f));
    e.setFuel((float)num_or(o, "fuel", 1.f));
    e.setDivergence((float)num_or(o, "divergence", 0.f));
    const std::pair<float, float> v = pair_or(o, "velocity", {0.f, 0.f});
    e.setVelocity(Vec3(v.first, v.second));
    e.setVelocityTransfer(bool_or(o, "transferVelocity", false));
    return e;
}

}  // namespace

const JsonValue& JsonValue::at(const std::string& key) const
{
    if (kind != OBJECT) bad("JSON: not an object (looking for '" + key + "')");
    auto it = obj.find(key);
    if (it == obj.end()) bad("JSON: missing key '" + key + "'");
    return it->second;
}
double JsonValue::number() const
{
    if (kind == NUMBER) return num;
    if (kind == BOOL) return b ? 1.0 : 0.0;
    bad("JSON: expected a number");
}
bool JsonValue::boolean() const
{
    if (kind == BOOL) return b;
    if (kind == NUMBER) return num != 0.0;
    bad("JSON: expected a boolean");
}
const std::string& JsonValue::string() const
{
    if (kind != STRING) bad("JSON: expected a string");
    return str;
}

JsonValue parseJson(const std::string& text) { return Parser(text).document(); }

SimulationMethod JsonSceneReader::simMethodFromName(const std::string& name)
{
    if (name == "fluid" || name == "flip") return SIMULATION_LIQUID;
    if (name == "smoke") return SIMULATION_SMOKE;
    if (name == "fire") return SIMULATION_FIRE;
    if (name == "nbflip") return SIMULATION_NBFLIP;
    return SIMULATION_LIQUID;
}

void JsonSceneReader::populateFlipSolverParamsFromJson(FlipSolverParameters* p, const JsonValue& s)
{
    const float scale = (float)num_or(s, "scale", 1.f);
    p->fluidDensity = (float)num_or(s, "density", 1.f);
    p->seed = (unsigned)num_or(s, "seed", 0);
    p->particlesPerCell = (int)s.at("particlesPerCell").number();
    const std::pair<float, float> g = pair_or(s, "globalAcceleration", {9.8f, 0.f});
    p->globalAcceleration = Vec3(g.first, g.second);
    p->resolution = (float)(int)s.at("resolution").number();
    p->fps = (int)s.at("fps").number();
    p->maxSubsteps = (int)num_or(s, "maxSubsteps", 30);
    p->picRatio = (float)num_or(s, "picRatio", 0.03);
    p->cflNumber = (float)num_or(s, "cflNumber", 10.f);
    p->particleScale = (float)num_or(s, "particleScale", 0.8);
    p->pcgIterLimit = (int)num_or(s, "pcgIterLimit", 200);
    p->viscosityEnabled = bool_or(s, "viscosityEnabled", false);
    p->domainSizeI = (float)s.at("domainSizeI").number() * scale;
    p->domainSizeJ = (float)s.at("domainSizeJ").number() * scale;
    p->useHeavyViscosity = bool_or(s, "heavyViscosity", false);
    p->sceneScale = scale;
    // the longer side gets `resolution` cells
    if (p->domainSizeI > p->domainSizeJ) {
        p->dx = static_cast<float>(p->domainSizeI) / p->resolution;
        p->gridSizeI = (int)p->resolution;
        p->gridSizeJ = (int)((static_cast<float>(p->domainSizeJ) / static_cast<float>(p->domainSizeI)) * p->resolution);
    } else {
        p->dx = static_cast<float>(p->domainSizeJ) / p->resolution;
        p->gridSizeJ = (int)p->resolution;
        p->gridSizeI = (int)((static_cast<float>(p->domainSizeI) / static_cast<float>(p->domainSizeJ)) * p->resolution);
    }
    const std::string h = s.has("parameterHandlingMethod") ? s.at("parameterHandlingMethod").string() : "particle";
    if (h == "particle") p->parameterHandlingMethod = PARTICLE;
    if (h == "hybrid") p->parameterHandlingMethod = HYBRID;
    if (h == "grid") p->parameterHandlingMethod = GRID;
    p->simulationMethod = simMethodFromName(s.at("simType").string());
}

SceneDescription JsonSceneReader::readSceneText(const std::string& text)
{
    const JsonValue doc = parseJson(text);
    SceneDescription d;
    populateFlipSolverParamsFromJson(&d.params, doc.at("settings"));
    const float scale = d.params.sceneScale;
    const JsonValue& objects = doc.at("solver").at("objects");
    if (objects.kind != JsonValue::ARRAY) bad("'objects' must be an array");
    // addObjectFromJson (jsonscenereader.cpp:255-293)
    for (const JsonValue& o : objects.arr) {
        if (!bool_or(o, "enabled", true)) continue;
        const std::string& type = o.at("type").string();
        if (type == "solid") {
            Geometry2d g = polygon(o, scale);
            // the reference reads friction as an int (tryGetValue with the
            // default 0, jsonscenereader.cpp:230): 0.3 becomes 0, 1.0 stays 1
            d.solids.emplace_back((float)(int)num_or(o, "friction", 0), g);
        } else if (type == "source") {
            d.sources.push_back(emitter(o, scale));
        } else if (type == "sink") {
            Geometry2d g = polygon(o, scale);
            d.sinks.emplace_back((float)num_or(o, "divergence", 0.f), g);
        } else if (type == "fluid") {
            d.fluids.push_back(emitter(o, scale));
        }
    }
    return d;
}

SceneDescription JsonSceneReader::readScene(const std::string& fileName)
{
    std::ifstream f(fileName);
    if (!f.is_open()) throw Error(FG_E_INVALID, "cannot open scene file " + fileName);
    std::stringstream ss;
    ss << f.rdbuf();
    return readSceneText(ss.str());
}

std::shared_ptr<FlipSolver> JsonSceneReader::loadJson(const std::string& fileName)
{
    try {
        SceneDescription d = readScene(fileName);
        if (d.params.simulationMethod == SIMULATION_SMOKE || d.params.simulationMethod == SIMULATION_FIRE)
            throw Error(FG_E_INVALID, "smoke and fire scenes are not supported (liquid solver only)");
        // jsonscenereader.cpp:50-62: an "nbflip" scene builds NBFlipSolver
        std::shared_ptr<FlipSolver> solver;
        if (d.params.simulationMethod == SIMULATION_NBFLIP) {
            NBFlipParameters np;
            static_cast<FlipSolverParameters&>(np) = d.params;
            solver = std::make_shared<NBFlipSolver>(&np);
        } else {
            solver = std::make_shared<FlipSolver>(&d.params);
        }
        solver->initAdditionalParameters();
        for (Obstacle& o : d.solids) solver->addGeometry(o);
        for (Emitter& e : d.sources) solver->addSource(e);
        for (Sink& s : d.sinks) solver->addSink(s);
        for (Emitter& e : d.fluids) solver->addInitialFluid(e);
        return solver;
    } catch (const std::exception& e) {
        std::cout << e.what() << std::endl;
        return std::shared_ptr<FlipSolver>();
    }
}

}  // namespace fghip
