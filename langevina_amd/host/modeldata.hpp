// modeldata.hpp -- the run parameters of ModelAData (ModelG/ModelA.h:32-269)
// read from the reference's input.json keys, and a small JSON reader that
// accepts what jsoncpp 1.9.5's default reader accepts in those files:
// // and /* */ comments and trailing commas (input.json:25 has one).
#pragma once

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace lgv_host {

// ------------------------------------------------------------------ JSON
struct Json {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.;
  bool integral = false;
  std::string str;
  std::vector<Json> arr;
  std::map<std::string, Json> obj;

  bool has(const std::string &k) const { return kind == Object && obj.count(k); }
  const Json &operator[](const std::string &k) const {
    static const Json null;
    auto it = obj.find(k);
    return it == obj.end() ? null : it->second;
  }
  // jsoncpp's asInt / asDouble / asBool / asString conversions
  int asInt() const {
    if (kind == Number) {
      if (num < -2147483648.0 || num > 2147483647.0) throw std::runtime_error("double out of Int range");
      return (int)num;
    }
    if (kind == Bool) return b ? 1 : 0;
    if (kind == Null) return 0;
    throw std::runtime_error("Value is not convertible to Int.");
  }
  double asDouble() const {
    if (kind == Number) return num;
    if (kind == Bool) return b ? 1.0 : 0.0;
    if (kind == Null) return 0.0;
    throw std::runtime_error("Value is not convertible to double.");
  }
  bool asBool() const {
    if (kind == Bool) return b;
    if (kind == Number) return num != 0.0 && !std::isnan(num);
    if (kind == Null) return false;
    throw std::runtime_error("Value is not convertible to bool.");
  }
  std::string asString() const {
    if (kind == String) return str;
    if (kind == Null) return "";
    if (kind == Bool) return b ? "true" : "false";
    if (kind == Number) {
      char buf[64];
      snprintf(buf, sizeof buf, integral ? "%.0f" : "%.17g", num);
      return buf;
    }
    throw std::runtime_error("Value is not convertible to string.");
  }
};

class JsonReader {
  const std::string &s;
  size_t i = 0;
  [[noreturn]] void err(const char *what) {
    throw std::runtime_error(std::string("JSON parse error at offset ") + std::to_string(i) +
                             ": " + what);
  }
  void ws() {
    for (;;) {
      while (i < s.size() && isspace((unsigned char)s[i])) i++;
      if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '/') {
        while (i < s.size() && s[i] != '\n') i++;
      } else if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '*') {
        size_t e = s.find("*/", i + 2);
        if (e == std::string::npos) err("unterminated comment");
        i = e + 2;
      } else {
        return;
      }
    }
  }
  std::string string_lit() {
    std::string out;
    i++;  // opening quote
    while (i < s.size() && s[i] != '"') {
      char c = s[i++];
      if (c == '\\') {
        if (i >= s.size()) err("bad escape");
        char e = s[i++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            unsigned v = std::stoul(s.substr(i, 4), nullptr, 16);
            i += 4;
            if (v < 0x80) out += (char)v;
            else out += '?';
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (i >= s.size()) err("unterminated string");
    i++;
    return out;
  }
  Json value() {
    ws();
    if (i >= s.size()) err("unexpected end");
    Json v;
    char c = s[i];
    if (c == '{') {
      v.kind = Json::Object;
      i++;
      for (;;) {
        ws();
        if (i < s.size() && s[i] == '}') { i++; break; }
        if (i >= s.size() || s[i] != '"') err("expected key");
        std::string k = string_lit();
        ws();
        if (i >= s.size() || s[i] != ':') err("expected ':'");
        i++;
        v.obj[k] = value();
        ws();
        if (i < s.size() && s[i] == ',') { i++; continue; }
        if (i < s.size() && s[i] == '}') { i++; break; }
        err("expected ',' or '}'");
      }
    } else if (c == '[') {
      v.kind = Json::Array;
      i++;
      for (;;) {
        ws();
        if (i < s.size() && s[i] == ']') { i++; break; }
        v.arr.push_back(value());
        ws();
        if (i < s.size() && s[i] == ',') { i++; continue; }
        if (i < s.size() && s[i] == ']') { i++; break; }
        err("expected ',' or ']'");
      }
    } else if (c == '"') {
      v.kind = Json::String;
      v.str = string_lit();
    } else if (s.compare(i, 4, "true") == 0) {
      v.kind = Json::Bool; v.b = true; i += 4;
    } else if (s.compare(i, 5, "false") == 0) {
      v.kind = Json::Bool; v.b = false; i += 5;
    } else if (s.compare(i, 4, "null") == 0) {
      i += 4;
    } else {
      size_t j = i;
      bool integral = true;
      if (s[j] == '-' || s[j] == '+') j++;
      while (j < s.size() && (isdigit((unsigned char)s[j]) || strchr(".eE+-", s[j]))) {
        if (strchr(".eE", s[j])) integral = false;
        j++;
      }
      if (j == i) err("unexpected character");
      v.kind = Json::Number;
      v.num = std::stod(s.substr(i, j - i));
      v.integral = integral;
      i = j;
    }
    return v;
  }

 public:
  explicit JsonReader(const std::string &text) : s(text) {}
  Json parse() {
    Json v = value();
    ws();
    if (i != s.size()) err("trailing characters");
    return v;
  }
};

// ------------------------------------------------------------------ params
// ModelATime, ModelA.h:32-68
struct ModelATime {
  double finaltime = 10., initialtime = 0., deltat = 0.24, time = 0.;
  double t() const { return time; }
  double dt() const { return deltat; }
  double tfinal() const { return finaltime; }
  void reset() { time = initialtime; }
  double adjust_finaltime(int saveFrequency) {
    finaltime = initialtime + deltat * saveFrequency *
                                  static_cast<int>((finaltime - initialtime) /
                                                   (deltat * saveFrequency));
    return finaltime;
  }
  void read(const Json &p) {
    if (p.has("finaltime")) finaltime = p["finaltime"].asDouble();
    if (p.has("initialtime")) initialtime = p["initialtime"].asDouble();
    if (p.has("deltat")) deltat = p["deltat"].asDouble();
    time = initialtime;
  }
  void print() const {
    printf("finaltime = %e\n", finaltime);
    printf("initialtime = %e\n", initialtime);
    printf("deltat  = %e\n", deltat);
  }
};

// ModelACoefficients, ModelA.h:71-111
struct ModelACoefficients {
  double mass0 = -4.70052, dmassdt = 0.;
  double lambda = 4., H = 0.003, chi = 5;
  double gamma = 1., diffusion = 0.3333333;
  double mass(double t) const { return mass0 + dmassdt * t; }
  double sigma() const { return diffusion * chi; }
  void read(const Json &p) {
    if (p.has("mass0")) mass0 = p["mass0"].asDouble();
    dmassdt = p.has("dmassdt") ? p["dmassdt"].asDouble() : 0.;
    if (p.has("lambda")) lambda = p["lambda"].asDouble();
    if (p.has("H")) H = p["H"].asDouble();
    if (p.has("chi")) chi = p["chi"].asDouble();
    if (p.has("gamma")) gamma = p["gamma"].asDouble();
    diffusion = p.has("diffusion") ? p["diffusion"].asDouble() : 1. / 3. * gamma;
  }
  void print() const {
    printf("mass0 = %e\n", mass0);
    printf("dmassdt = %e\n", dmassdt);
    printf("lambda = %e\n", lambda);
    printf("H = %e\n", H);
    printf("chi = %e\n", chi);
    printf("gamma = %e\n", gamma);
    printf("diffusion = %e\n", diffusion);
  }
};

// ModelAHandlerData, ModelA.h:114-181
struct ModelAHandlerData {
  std::string evolverType = "PV2HBSplit23";
  int seed = 10;
  bool restart = false;
  std::string outputfiletag = "o4output";
  int saveFrequency = 3;
  int writeFrequency = -1;
  bool eventmode = false;
  int nevents = 1;
  int current_event = 0;
  double thermalization_time = 0.;
  bool quench_mode = false;
  double quench_mode_mass0 = -4.70052;
  std::string precision = "fp64";  // build extension: "fp32" stores FP32 fields
  void read(const Json &p) {
    if (p.has("evolverType")) evolverType = p["evolverType"].asString();
    if (p.has("seed")) seed = p["seed"].asInt();
    restart = p.has("restart") ? p["restart"].asBool() : false;
    outputfiletag = p.has("outputfiletag") ? p["outputfiletag"].asString() : "o4output";
    if (p.has("saveFrequency")) saveFrequency = p["saveFrequency"].asInt();
    if (p.has("writeFrequency")) writeFrequency = p["writeFrequency"].asInt();
    if (p.has("thermalization_time")) thermalization_time = p["thermalization_time"].asDouble();
    if (p.has("quench_mode")) quench_mode = p["quench_mode"].asBool();
    if (p.has("quench_mode_mass0")) quench_mode_mass0 = p["quench_mode_mass0"].asDouble();
    if (p.has("eventmode")) eventmode = p["eventmode"].asBool();
    if (p.has("nevents")) nevents = p["nevents"].asInt();
    if (p.has("precision")) precision = p["precision"].asString();
  }
  void print() const {
    printf("evolverType = %s\n", evolverType.c_str());
    printf("seed = %d\n", seed);
    printf("restart = %s\n", restart ? "true" : "false");
    printf("outputfiletag = %s\n", outputfiletag.c_str());
    printf("saveFrequency = %d\n", saveFrequency);
    printf("thermalization_time = %e\n", thermalization_time);
    printf("quench_mode = %s\n", quench_mode ? "true" : "false");
    printf("quench_mode_mass0 = %e\n", quench_mode_mass0);
    printf("eventmode = %s\n", eventmode ? "true" : "false");
    printf("nevents = %d\n", nevents);
  }
};

// ModelAData, ModelA.h:185-269 (NY = NZ = NX; LX read as an integer)
struct ModelAData {
  int NX = 16, NY = 16, NZ = 16;
  double LX = 16., LY = 16., LZ = 16.;
  ModelATime atime;
  ModelACoefficients acoefficients;
  ModelAHandlerData ahandler;
  double hX() const { return LX / NX; }
  double mass() const { return acoefficients.mass(atime.t()); }
  explicit ModelAData(Json &params) {
    if (!params.has("NX")) throw std::runtime_error("input: NX is required");
    NX = params["NX"].asInt();
    NY = NX;
    NZ = NX;
    LX = params.has("LX") ? params["LX"].asInt() : NX;
    LY = LX;
    LZ = LX;
    atime.read(params);
    acoefficients.read(params);
    ahandler.read(params);
    const double finaltime = atime.adjust_finaltime(ahandler.saveFrequency);
    Json f;
    f.kind = Json::Number;
    f.num = finaltime;
    params.obj["finaltime"] = f;
    print();
  }
  void print() const {
    printf("NX = %d\n", NX);
    printf("NY = %d\n", NY);
    printf("NZ = %d\n", NZ);
    printf("LX = %e\n", LX);
    printf("LY = %e\n", LY);
    printf("LZ = %e\n", LZ);
    atime.print();
    acoefficients.print();
    ahandler.print();
  }
};

}  // namespace lgv_host
