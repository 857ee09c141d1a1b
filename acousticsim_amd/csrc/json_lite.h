// Minimal JSON reader for the cloth configuration (objects, arrays, numbers,
// strings, booleans, null).  The reference reads the same documents with
// jsoncpp through cJsonUtil::ParseAs* (src/utils/JsonUtil.cpp:139-200); a
// missing key there exits the process, here it throws std::runtime_error that
// the C ABI turns into an error status.
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace bw
{

struct Json
{
    enum Kind
    {
        Null,
        Bool,
        Number,
        String,
        Array,
        Object
    } kind = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    static Json parse(const std::string &text);

    bool has(const std::string &k) const { return kind == Object && obj.count(k); }
    const Json &at(const std::string &k) const
    {
        auto it = obj.find(k);
        if (kind != Object || it == obj.end())
            throw std::runtime_error("config: missing key \"" + k + "\"");
        return it->second;
    }
    double as_double() const
    {
        if (kind == Number)
            return num;
        if (kind == Bool)
            return b ? 1.0 : 0.0;
        throw std::runtime_error("config: expected a number");
    }
    int as_int() const { return static_cast<int>(as_double()); }
    bool as_bool() const
    {
        if (kind == Bool)
            return b;
        if (kind == Number)
            return num != 0;
        throw std::runtime_error("config: expected a boolean");
    }
    const std::string &as_string() const
    {
        if (kind != String)
            throw std::runtime_error("config: expected a string");
        return str;
    }
    std::vector<double> as_vector() const
    {
        if (kind != Array)
            throw std::runtime_error("config: expected an array");
        std::vector<double> v;
        for (auto &e : arr)
            v.push_back(e.as_double());
        return v;
    }
    double get_double(const std::string &k, double dflt) const
    {
        return has(k) ? at(k).as_double() : dflt;
    }
    std::string get_string(const std::string &k, const std::string &dflt) const
    {
        return has(k) ? at(k).as_string() : dflt;
    }
};

} // namespace bw
