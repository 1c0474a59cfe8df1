// A small ordered JSON document writer whose output matches nlohmann::ordered_json dumped with
// std::setw(4) (the reference writes metadata.json and statistics.json that way,
// potree_exporter.cu:170-172, json_exporter.h:96-98): 4-space indentation, one array element per
// line, integers as integers and doubles in the shortest round-trip form with nlohmann's
// fixed/exponential rules (an integral double keeps a trailing ".0").
#pragma once

#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace oc {

class Json
{
public:
    enum Kind
    {
        NUL,
        OBJECT,
        ARRAY,
        STRING,
        INT,
        UINT,
        DOUBLE,
        BOOL
    };

    Json() : itsKind(OBJECT) {}
    static Json object() { return Json(OBJECT); }
    static Json array() { return Json(ARRAY); }
    static Json str(const std::string& s)
    {
        Json j(STRING);
        j.itsStr = s;
        return j;
    }
    static Json i(int64_t v)
    {
        Json j(INT);
        j.itsInt = v;
        return j;
    }
    static Json u(uint64_t v)
    {
        Json j(UINT);
        j.itsUint = v;
        return j;
    }
    static Json d(double v)
    {
        Json j(DOUBLE);
        j.itsDouble = v;
        return j;
    }
    static Json b(bool v)
    {
        Json j(BOOL);
        j.itsBool = v;
        return j;
    }

    // object member (inserted in order; an existing key is replaced in place)
    Json& operator[](const std::string& key)
    {
        itsKind = OBJECT;
        for (auto& kv : itsMembers)
            if (kv.first == key)
                return *kv.second;
        itsMembers.emplace_back(key, std::unique_ptr<Json>(new Json(NUL)));
        return *itsMembers.back().second;
    }

    Json& push(Json v)
    {
        itsKind = ARRAY;
        itsItems.push_back(std::unique_ptr<Json>(new Json(std::move(v))));
        return *itsItems.back();
    }

    Json(Json&&) noexcept = default;
    Json& operator=(Json&&) noexcept = default;

    std::string dump(int indent = 4) const
    {
        std::string out;
        write(out, indent, 0);
        return out;
    }

    static std::string formatDouble(double v)
    {
        if (v != v)
            return "null";
        if (v == 0.0)
            return std::signbit(v) ? "-0.0" : "0.0";
        char buf[64];
        auto res = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
        std::string sci(buf, res.ptr);
        bool neg = sci[0] == '-';
        if (neg)
            sci.erase(0, 1);
        size_t epos      = sci.find('e');
        std::string mant = sci.substr(0, epos);
        int exp10        = std::stoi(sci.substr(epos + 1));
        std::string digits;
        for (char c : mant)
            if (c != '.')
                digits.push_back(c);
        int k = static_cast<int>(digits.size());
        int n = exp10 + 1;  // value = 0.d1d2...dk * 10^n
        std::string s;
        if (k <= n && n <= 15)
            s = digits + std::string(n - k, '0') + ".0";
        else if (0 < n && n <= 15)
            s = digits.substr(0, n) + "." + digits.substr(n);
        else if (-4 < n && n <= 0)
            s = "0." + std::string(-n, '0') + digits;
        else
        {
            s = digits.substr(0, 1);
            if (k > 1)
                s += "." + digits.substr(1);
            int e = n - 1;
            char eb[16];
            std::snprintf(eb, sizeof(eb), "e%c%02d", e < 0 ? '-' : '+', e < 0 ? -e : e);
            s += eb;
        }
        return neg ? "-" + s : s;
    }

private:
    explicit Json(Kind k) : itsKind(k) {}

    static std::string escape(const std::string& s)
    {
        std::string o = "\"";
        for (char c : s)
        {
            switch (c)
            {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\n': o += "\\n"; break;
            case '\t': o += "\\t"; break;
            default: o.push_back(c);
            }
        }
        return o + "\"";
    }

    void write(std::string& out, int indent, int depth) const
    {
        std::string pad(static_cast<size_t>(indent) * (depth + 1), ' ');
        std::string close(static_cast<size_t>(indent) * depth, ' ');
        switch (itsKind)
        {
        case NUL: out += "null"; break;
        case STRING: out += escape(itsStr); break;
        case INT: out += std::to_string(itsInt); break;
        case UINT: out += std::to_string(itsUint); break;
        case DOUBLE: out += formatDouble(itsDouble); break;
        case BOOL: out += itsBool ? "true" : "false"; break;
        case OBJECT:
            if (itsMembers.empty())
            {
                out += "{}";
                break;
            }
            out += "{\n";
            for (size_t i = 0; i < itsMembers.size(); ++i)
            {
                out += pad + escape(itsMembers[i].first) + ": ";
                itsMembers[i].second->write(out, indent, depth + 1);
                out += (i + 1 < itsMembers.size()) ? ",\n" : "\n";
            }
            out += close + "}";
            break;
        case ARRAY:
            if (itsItems.empty())
            {
                out += "[]";
                break;
            }
            out += "[\n";
            for (size_t i = 0; i < itsItems.size(); ++i)
            {
                out += pad;
                itsItems[i]->write(out, indent, depth + 1);
                out += (i + 1 < itsItems.size()) ? ",\n" : "\n";
            }
            out += close + "]";
            break;
        }
    }

    Kind itsKind;
    std::string itsStr;
    int64_t itsInt    = 0;
    uint64_t itsUint  = 0;
    double itsDouble  = 0.0;
    bool itsBool      = false;
    std::vector<std::pair<std::string, std::unique_ptr<Json>>> itsMembers;
    std::vector<std::unique_ptr<Json>> itsItems;
};

}  // namespace oc
