/*
 * mcx_json.h -- a small JSON document model (parse, query, print) for the
 * input files and the .jnii/.jdat writers.  Object member order is kept,
 * as the Shapes section is an ordered list of single-key objects.
 */
#ifndef MCX_AMD_JSON_H
#define MCX_AMD_JSON_H

#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace mcx {

struct Json {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;

    static Json parse(const std::string& text);  /* throws mcx::Error on bad input */
    std::string dump(int indent = 1) const;

    bool is_obj() const { return type == Object; }
    bool is_arr() const { return type == Array; }
    bool is_num() const { return type == Number || type == Bool; }
    bool is_str() const { return type == String; }
    const Json* get(const char* key) const;  /* object member or nullptr */
    size_t size() const { return type == Array ? arr.size() : (type == Object ? obj.size() : 0); }
    double as_num(double dflt = 0.0) const { return type == Number ? num : (type == Bool ? (b ? 1.0 : 0.0) : dflt); }

    /* builders */
    static Json number(double v) { Json j; j.type = Number; j.num = v; return j; }
    static Json string(std::string s) { Json j; j.type = String; j.str = std::move(s); return j; }
    static Json array() { Json j; j.type = Array; return j; }
    static Json object() { Json j; j.type = Object; return j; }
    Json& set(const std::string& k, Json v) { obj.emplace_back(k, std::move(v)); return obj.back().second; }
    Json& push(Json v) { arr.push_back(std::move(v)); return arr.back(); }
};

}  // namespace mcx

#endif
