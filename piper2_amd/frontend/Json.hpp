// A small JSON document model and parser for scene descriptions (RFC 8259: objects, arrays,
// strings with escapes, numbers, true/false/null). Objects keep their keys in file order. Errors
// throw std::runtime_error with the line and column of the offending character.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace PiperFrontend {

class Json {
public:
    enum class Type { Null, Bool, Number, String, Array, Object };

    Type type() const noexcept { return mType; }
    bool isNull() const noexcept { return mType == Type::Null; }
    bool isNumber() const noexcept { return mType == Type::Number; }
    bool isString() const noexcept { return mType == Type::String; }
    bool isArray() const noexcept { return mType == Type::Array; }
    bool isObject() const noexcept { return mType == Type::Object; }

    double number() const {
        expect(Type::Number, "a number");
        return mNumber;
    }
    bool boolean() const {
        expect(Type::Bool, "true or false");
        return mBool;
    }
    const std::string& string() const {
        expect(Type::String, "a string");
        return mString;
    }
    const std::vector<Json>& array() const {
        expect(Type::Array, "an array");
        return mArray;
    }
    const std::vector<std::pair<std::string, Json>>& members() const {
        expect(Type::Object, "an object");
        return mMembers;
    }
    // the member `key` of an object, or nullptr
    const Json* find(const std::string& key) const {
        expect(Type::Object, "an object");
        for (const auto& [k, v] : mMembers)
            if (k == key) return &v;
        return nullptr;
    }
    // the member `key`; throws naming the key when it is missing
    const Json& at(const std::string& key) const {
        const Json* v = find(key);
        if (!v) throw std::runtime_error("missing key \"" + key + "\"" + where());
        return *v;
    }
    // where this value starts in its file, for messages
    std::string where() const { return mLine ? " (line " + std::to_string(mLine) + ")" : std::string(); }

    static Json parse(const std::string& text);

    // values built in code (component configs made by factories)
    static Json makeNumber(double v) {
        Json j;
        j.mType = Type::Number, j.mNumber = v;
        return j;
    }
    static Json makeString(std::string v) {
        Json j;
        j.mType = Type::String, j.mString = std::move(v);
        return j;
    }
    static Json makeArray(std::vector<Json> v) {
        Json j;
        j.mType = Type::Array, j.mArray = std::move(v);
        return j;
    }
    static Json makeObject(std::vector<std::pair<std::string, Json>> members) {
        Json j;
        j.mType = Type::Object, j.mMembers = std::move(members);
        return j;
    }
    // the same object with member `key` replaced (or added)
    Json with(const std::string& key, Json v) const {
        expect(Type::Object, "an object");
        Json j = *this;
        for (auto& [k, old] : j.mMembers)
            if (k == key) {
                old = std::move(v);
                return j;
            }
        j.mMembers.emplace_back(key, std::move(v));
        return j;
    }

private:
    friend class JsonParser;
    void expect(Type t, const char* what) const {
        if (mType != t) throw std::runtime_error(std::string("expected ") + what + where());
    }
    Type mType = Type::Null;
    bool mBool = false;
    double mNumber = 0.0;
    std::string mString;
    std::vector<Json> mArray;
    std::vector<std::pair<std::string, Json>> mMembers;
    uint32_t mLine = 0;
};

class JsonParser {
public:
    explicit JsonParser(const std::string& text) : s(text) {}
    Json document() {
        Json v = value(0);
        skip();
        if (i != s.size()) fail("trailing characters after the document");
        return v;
    }

private:
    const std::string& s;
    size_t i = 0;
    uint32_t line = 1;
    size_t lineStart = 0;

    [[noreturn]] void fail(const std::string& msg) const {
        throw std::runtime_error("JSON: " + msg + " at line " + std::to_string(line) + ", column " +
                                 std::to_string(i - lineStart + 1));
    }
    void skip() {
        while (i < s.size()) {
            const char c = s[i];
            if (c == '\n') {
                ++line;
                lineStart = i + 1;
            } else if (c != ' ' && c != '\t' && c != '\r') {
                break;
            }
            ++i;
        }
    }
    bool literal(const char* word) {
        size_t n = 0;
        while (word[n]) ++n;
        if (s.compare(i, n, word) != 0) return false;
        i += n;
        return true;
    }
    Json value(int depth) {
        if (depth > 256) fail("nesting deeper than 256");
        skip();
        if (i >= s.size()) fail("unexpected end of input");
        Json v;
        v.mLine = line;
        const char c = s[i];
        if (c == '{') {
            v.mType = Json::Type::Object;
            ++i;
            skip();
            if (i < s.size() && s[i] == '}') {
                ++i;
                return v;
            }
            while (true) {
                skip();
                if (i >= s.size() || s[i] != '"') fail("expected a string key");
                std::string key = str();
                skip();
                if (i >= s.size() || s[i] != ':') fail("expected ':'");
                ++i;
                v.mMembers.emplace_back(std::move(key), value(depth + 1));
                skip();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == '}') {
                    ++i;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.mType = Json::Type::Array;
            ++i;
            skip();
            if (i < s.size() && s[i] == ']') {
                ++i;
                return v;
            }
            while (true) {
                v.mArray.push_back(value(depth + 1));
                skip();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == ']') {
                    ++i;
                    return v;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.mType = Json::Type::String;
            v.mString = str();
            return v;
        }
        if (literal("true")) {
            v.mType = Json::Type::Bool;
            v.mBool = true;
            return v;
        }
        if (literal("false")) {
            v.mType = Json::Type::Bool;
            return v;
        }
        if (literal("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) {
            v.mType = Json::Type::Number;
            v.mNumber = number();
            return v;
        }
        fail(std::string("unexpected character '") + c + "'");
    }
    double number() {
        const size_t b = i;
        if (s[i] == '-') ++i;
        auto digits = [&] {
            const size_t d = i;
            while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
            if (i == d) fail("malformed number");
        };
        if (i < s.size() && s[i] == '0') ++i;
        else digits();
        if (i < s.size() && s[i] == '.') {
            ++i;
            digits();
        }
        if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
            ++i;
            if (i < s.size() && (s[i] == '+' || s[i] == '-')) ++i;
            digits();
        }
        return std::strtod(s.substr(b, i - b).c_str(), nullptr);
    }
    static void utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += static_cast<char>(cp);
        } else if (cp < 0x800) {
            out += static_cast<char>(0xC0 | (cp >> 6));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += static_cast<char>(0xE0 | (cp >> 12));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else {
            out += static_cast<char>(0xF0 | (cp >> 18));
            out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (i + 4 > s.size()) fail("short \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            const char h = s[i++];
            v <<= 4;
            if (h >= '0' && h <= '9') v |= static_cast<uint32_t>(h - '0');
            else if (h >= 'a' && h <= 'f') v |= static_cast<uint32_t>(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= static_cast<uint32_t>(h - 'A' + 10);
            else fail("bad hex digit in \\u escape");
        }
        return v;
    }
    std::string str() {
        ++i;  // opening quote
        std::string out;
        while (true) {
            if (i >= s.size()) fail("unterminated string");
            const char c = s[i++];
            if (c == '"') return out;
            if (static_cast<unsigned char>(c) < 0x20) fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i >= s.size()) fail("unterminated escape");
            const char e = s[i++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00) {  // surrogate pair
                        if (i + 2 > s.size() || s[i] != '\\' || s[i + 1] != 'u') fail("lone high surrogate");
                        i += 2;
                        const uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo >= 0xE000) fail("bad low surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp < 0xE000) {
                        fail("lone low surrogate");
                    }
                    utf8(out, cp);
                    break;
                }
                default: fail(std::string("bad escape \\") + e);
            }
        }
    }
};

inline Json Json::parse(const std::string& text) { return JsonParser(text).document(); }

}  // namespace PiperFrontend
