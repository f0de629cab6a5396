// json_scene.h -- scene files of the reference's JSON format, read into the
// reference-named C++ API (JsonSceneReader, Utils/jsonscenereader.h), with
// a small self-contained JSON parser (the reference uses nlohmann::json).
//
// Format (README.md:34-135): {"settings": {...}, "solver": {"objects":
// [...]}}; settings give the domain size (units), resolution (cells along the
// shorter side), fps, density, particles per cell, gravity and the solver
// knobs; objects are polygons of type solid / source / sink / fluid with
// their own attributes.  Coordinates: (x, I, U) points down, (y, J, V) right.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "flipsolver2d_hip.h"

namespace fghip {

// A parsed JSON value.
struct JsonValue {
    enum Kind { NUL, BOOL, NUMBER, STRING, ARRAY, OBJECT } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JsonValue> arr;
    std::map<std::string, JsonValue> obj;

    bool has(const std::string& key) const { return kind == OBJECT && obj.count(key) != 0; }
    const JsonValue& at(const std::string& key) const;   // throws Error if missing
    double number() const;
    bool boolean() const;
    const std::string& string() const;
};

// Parse JSON text; throws Error(FG_E_INVALID) with the offset of a syntax error.
JsonValue parseJson(const std::string& text);

// Everything a scene file describes, before a solver is built from it.
struct SceneDescription {
    FlipSolverParameters params;
    std::vector<Obstacle> solids;
    std::vector<Emitter> sources;
    std::vector<Sink> sinks;
    std::vector<Emitter> fluids;
};

class JsonSceneReader {
public:
    // The reference's entry point (jsonscenereader.cpp:8-77): a solver with
    // the scene's parameters and objects, or an empty pointer (with the
    // error printed) when the file cannot be read.  "flip" scenes build the
    // FLIP liquid solver, "nbflip" ones NBFlipSolver; smoke and fire scenes
    // are refused.
    static std::shared_ptr<FlipSolver> loadJson(const std::string& fileName);
    // The same file as a description (throws Error).
    static SceneDescription readScene(const std::string& fileName);
    static SceneDescription readSceneText(const std::string& text);
    // FlipSolverParameters from the settings object (jsonscenereader.cpp:79-136).
    static void populateFlipSolverParamsFromJson(FlipSolverParameters* p, const JsonValue& settings);
    static SimulationMethod simMethodFromName(const std::string& name);
};

}  // namespace fghip
