#include "json_lite.h"

#include <cctype>
#include <cstdlib>

namespace bw
{
namespace
{
struct Parser
{
    const std::string &s;
    size_t i = 0;

    explicit Parser(const std::string &t) : s(t) {}

    [[noreturn]] void fail(const char *what)
    {
        throw std::runtime_error(std::string("config: JSON parse error (") + what +
                                 ") at offset " + std::to_string(i));
    }
    void ws()
    {
        while (i < s.size())
        {
            if (std::isspace(static_cast<unsigned char>(s[i])))
                i++;
            else if (s.compare(i, 2, "//") == 0)
                while (i < s.size() && s[i] != '\n')
                    i++;
            else
                break;
        }
    }
    bool lit(const char *w)
    {
        size_t n = std::char_traits<char>::length(w);
        if (s.compare(i, n, w) == 0)
        {
            i += n;
            return true;
        }
        return false;
    }
    std::string string_body()
    {
        std::string out;
        i++; // opening quote
        while (i < s.size() && s[i] != '"')
        {
            char c = s[i++];
            if (c == '\\')
            {
                if (i >= s.size())
                    fail("escape");
                char e = s[i++];
                switch (e)
                {
                case 'n': out.push_back('\n'); break;
                case 't': out.push_back('\t'); break;
                case 'r': out.push_back('\r'); break;
                case 'b': out.push_back('\b'); break;
                case 'f': out.push_back('\f'); break;
                case 'u':
                    if (i + 4 > s.size())
                        fail("unicode escape");
                    out.push_back('?');
                    i += 4;
                    break;
                default: out.push_back(e);
                }
            }
            else
                out.push_back(c);
        }
        if (i >= s.size())
            fail("unterminated string");
        i++;
        return out;
    }
    Json value()
    {
        ws();
        if (i >= s.size())
            fail("unexpected end");
        Json j;
        char c = s[i];
        if (c == '{')
        {
            j.kind = Json::Object;
            i++;
            ws();
            if (i < s.size() && s[i] == '}')
            {
                i++;
                return j;
            }
            for (;;)
            {
                ws();
                if (i >= s.size() || s[i] != '"')
                    fail("object key");
                std::string k = string_body();
                ws();
                if (i >= s.size() || s[i] != ':')
                    fail("colon");
                i++;
                j.obj[k] = value();
                ws();
                if (i < s.size() && s[i] == ',')
                {
                    i++;
                    continue;
                }
                if (i < s.size() && s[i] == '}')
                {
                    i++;
                    return j;
                }
                fail("object");
            }
        }
        if (c == '[')
        {
            j.kind = Json::Array;
            i++;
            ws();
            if (i < s.size() && s[i] == ']')
            {
                i++;
                return j;
            }
            for (;;)
            {
                j.arr.push_back(value());
                ws();
                if (i < s.size() && s[i] == ',')
                {
                    i++;
                    continue;
                }
                if (i < s.size() && s[i] == ']')
                {
                    i++;
                    return j;
                }
                fail("array");
            }
        }
        if (c == '"')
        {
            j.kind = Json::String;
            j.str = string_body();
            return j;
        }
        if (lit("true"))
        {
            j.kind = Json::Bool;
            j.b = true;
            return j;
        }
        if (lit("false"))
        {
            j.kind = Json::Bool;
            j.b = false;
            return j;
        }
        if (lit("null"))
            return j;
        const char *begin = s.c_str() + i;
        char *end = nullptr;
        double v = std::strtod(begin, &end);
        if (end == begin)
            fail("value");
        i += static_cast<size_t>(end - begin);
        j.kind = Json::Number;
        j.num = v;
        return j;
    }
};
} // namespace

Json Json::parse(const std::string &text)
{
    Parser p(text);
    Json j = p.value();
    p.ws();
    if (p.i != text.size())
        p.fail("trailing characters");
    return j;
}

} // namespace bw
