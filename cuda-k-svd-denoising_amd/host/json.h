// Minimal JSON reader with the accessor semantics the reference's config
// parser relies on (jsonxx, src/lib/jsonLib/jsonxx.h:393-412):
//   get<T>(key)           value of type T; a missing key or a wrong type is an
//                         error (jsonxx asserts; here a json::Error is thrown);
//   get<T>(key, fallback) the value, or the fallback when missing / wrong type.
// Numbers are doubles (jsonxx::Number).
#pragma once

#include <istream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace json {

typedef double Number;
typedef std::string String;

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Object;
class Array;

class Value {
public:
    enum Kind { NUL, BOOL, NUMBER, STRING, ARRAY, OBJECT };
    Kind kind = NUL;
    bool b = false;
    double num = 0;
    std::string str;
    std::shared_ptr<Array> arr;
    std::shared_ptr<Object> obj;
};

class Array {
public:
    std::vector<Value> items;
    size_t size() const { return items.size(); }
    template <typename T> T get(size_t i) const;
};

class Object {
public:
    std::map<std::string, Value> kv;
    bool parse(std::istream &in);  // false (with a message in last_error) on malformed input
    bool parse(const std::string &text);
    std::string last_error;
    template <typename T> bool has(const std::string &key) const;
    template <typename T> T get(const std::string &key) const;
    template <typename T> T get(const std::string &key, const T &fallback) const;
};

namespace detail {
template <typename T> bool is(const Value &v);
template <typename T> T as(const Value &v);
template <> bool is<double>(const Value &v);
template <> bool is<bool>(const Value &v);
template <> bool is<std::string>(const Value &v);
template <> bool is<Array>(const Value &v);
template <> bool is<Object>(const Value &v);
template <> double as<double>(const Value &v);
template <> bool as<bool>(const Value &v);
template <> std::string as<std::string>(const Value &v);
template <> Array as<Array>(const Value &v);
template <> Object as<Object>(const Value &v);
}  // namespace detail

template <typename T> bool Object::has(const std::string &key) const
{
    auto it = kv.find(key);
    return it != kv.end() && detail::is<T>(it->second);
}

template <typename T> T Object::get(const std::string &key) const
{
    auto it = kv.find(key);
    if (it == kv.end()) throw Error("missing required key \"" + key + "\"");
    if (!detail::is<T>(it->second)) throw Error("key \"" + key + "\" has the wrong type");
    return detail::as<T>(it->second);
}

template <typename T> T Object::get(const std::string &key, const T &fallback) const
{
    auto it = kv.find(key);
    if (it == kv.end() || !detail::is<T>(it->second)) return fallback;
    return detail::as<T>(it->second);
}

template <typename T> T Array::get(size_t i) const
{
    if (i >= items.size()) throw Error("array index out of range");
    if (!detail::is<T>(items[i])) throw Error("array item has the wrong type");
    return detail::as<T>(items[i]);
}

}  // namespace json
