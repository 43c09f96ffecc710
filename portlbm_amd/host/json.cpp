// JSON document model and the settings-file mapping.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <sstream>

#include "lbm/exceptions/exceptions.hpp"
#include "lbm/file_interaction/file_interaction.hpp"
#include "lbm/file_interaction/json.hpp"

namespace lbm
{
namespace file_interaction
{

using exceptions::json::MissingPropertyException;
using exceptions::json::UnknownPropertyException;

// ---------------------------------------------------------------------------
// JsonValue
// ---------------------------------------------------------------------------
JsonValue JsonValue::boolean(bool b)
{
    JsonValue v;
    v.kind_ = Kind::Bool;
    v.bool_ = b;
    return v;
}

JsonValue JsonValue::number(double d)
{
    JsonValue v;
    v.kind_ = Kind::Number;
    v.number_ = d;
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", d);
    // shortest text that reads back to the same double
    for (int prec = 1; prec <= 17; ++prec)
    {
        char b2[40];
        std::snprintf(b2, sizeof b2, "%.*g", prec, d);
        if (std::strtod(b2, nullptr) == d)
        {
            std::snprintf(buf, sizeof buf, "%s", b2);
            break;
        }
    }
    v.text_ = buf;
    if (v.text_.find_first_of(".eEn") == std::string::npos)
        v.text_ += ".0";  // keep floats recognisable as such
    return v;
}

JsonValue JsonValue::unsigned_number(std::uint64_t u)
{
    JsonValue v;
    v.kind_ = Kind::Number;
    v.number_ = static_cast<double>(u);
    v.text_ = std::to_string(u);
    return v;
}

JsonValue JsonValue::string(std::string s)
{
    JsonValue v;
    v.kind_ = Kind::String;
    v.text_ = std::move(s);
    return v;
}

JsonValue JsonValue::array()
{
    JsonValue v;
    v.kind_ = Kind::Array;
    return v;
}

JsonValue JsonValue::object()
{
    JsonValue v;
    v.kind_ = Kind::Object;
    return v;
}

bool JsonValue::as_bool(const std::string &what) const
{
    if (kind_ != Kind::Bool)
        throw UnknownPropertyException("JSON property \"" + what + "\" must be a boolean");
    return bool_;
}

double JsonValue::as_double(const std::string &what) const
{
    if (kind_ != Kind::Number)
        throw UnknownPropertyException("JSON property \"" + what + "\" must be a number");
    return number_;
}

std::uint64_t JsonValue::as_unsigned(const std::string &what) const
{
    if (kind_ != Kind::Number)
        throw UnknownPropertyException("JSON property \"" + what + "\" must be a number");
    if (text_.find_first_of(".eE-") == std::string::npos)
    {
        errno = 0;
        const unsigned long long u = std::strtoull(text_.c_str(), nullptr, 10);
        if (errno == 0)
            return u;
    }
    if (number_ < 0 || number_ != std::floor(number_) || number_ > 1.8446744073709552e19)
        throw UnknownPropertyException("JSON property \"" + what + "\" must be a non-negative integer");
    return static_cast<std::uint64_t>(number_);
}

const std::string &JsonValue::as_string(const std::string &what) const
{
    if (kind_ != Kind::String)
        throw UnknownPropertyException("JSON property \"" + what + "\" must be a string");
    return text_;
}

const JsonValue *JsonValue::find(const std::string &key) const
{
    if (kind_ != Kind::Object)
        return nullptr;
    for (const auto &kv : object_)
        if (kv.first == key)
            return &kv.second;
    return nullptr;
}

const JsonValue &JsonValue::at(const std::string &key) const
{
    const JsonValue *v = find(key);
    if (!v)
        throw MissingPropertyException("JSON property \"" + key + "\" is missing");
    return *v;
}

JsonValue &JsonValue::operator[](const std::string &key)
{
    if (kind_ == Kind::Null)
        kind_ = Kind::Object;
    for (auto &kv : object_)
        if (kv.first == key)
            return kv.second;
    object_.emplace_back(key, JsonValue());
    return object_.back().second;
}

void JsonValue::push_back(JsonValue v)
{
    if (kind_ == Kind::Null)
        kind_ = Kind::Array;
    array_.push_back(std::move(v));
}

namespace
{
void escape_to(std::string &out, const std::string &s)
{
    out += '"';
    for (unsigned char c : s)
    {
        switch (c)
        {
        case '"': out += "\\\""; break;
        case '\\': out += "\\\\"; break;
        case '\n': out += "\\n"; break;
        case '\t': out += "\\t"; break;
        case '\r': out += "\\r"; break;
        case '\b': out += "\\b"; break;
        case '\f': out += "\\f"; break;
        default:
            if (c < 0x20)
            {
                char buf[8];
                std::snprintf(buf, sizeof buf, "\\u%04x", c);
                out += buf;
            }
            else
                out += static_cast<char>(c);
        }
    }
    out += '"';
}
}  // namespace

void JsonValue::dump_to(std::string &out, int indent, int level) const
{
    const std::string pad(static_cast<std::size_t>(indent * (level + 1)), ' ');
    const std::string pad_close(static_cast<std::size_t>(indent * level), ' ');
    switch (kind_)
    {
    case Kind::Null: out += "null"; break;
    case Kind::Bool: out += bool_ ? "true" : "false"; break;
    case Kind::Number: out += text_; break;
    case Kind::String: escape_to(out, text_); break;
    case Kind::Array:
        if (array_.empty())
        {
            out += "[]";
            break;
        }
        out += "[\n";
        for (std::size_t i = 0; i < array_.size(); ++i)
        {
            out += pad;
            array_[i].dump_to(out, indent, level + 1);
            out += i + 1 < array_.size() ? ",\n" : "\n";
        }
        out += pad_close + "]";
        break;
    case Kind::Object:
        if (object_.empty())
        {
            out += "{}";
            break;
        }
        out += "{\n";
        for (std::size_t i = 0; i < object_.size(); ++i)
        {
            out += pad;
            escape_to(out, object_[i].first);
            out += ": ";
            object_[i].second.dump_to(out, indent, level + 1);
            out += i + 1 < object_.size() ? ",\n" : "\n";
        }
        out += pad_close + "}";
        break;
    }
}

std::string JsonValue::dump(int indent) const
{
    std::string out;
    dump_to(out, indent, 0);
    return out;
}

namespace
{
class Parser
{
  public:
    explicit Parser(const std::string &text) : s_(text) {}

    JsonValue document()
    {
        JsonValue v = value();
        skip_ws();
        if (pos_ != s_.size())
            error("trailing characters");
        return v;
    }

  private:
    [[noreturn]] void error(const std::string &what) const
    {
        throw UnknownPropertyException("JSON syntax error at offset " + std::to_string(pos_) + ": " + what);
    }

    void skip_ws()
    {
        while (pos_ < s_.size() && (s_[pos_] == ' ' || s_[pos_] == '\n' || s_[pos_] == '\t' || s_[pos_] == '\r'))
            ++pos_;
    }

    bool consume(const char *word)
    {
        std::size_t n = 0;
        while (word[n])
            ++n;
        if (s_.compare(pos_, n, word) == 0)
        {
            pos_ += n;
            return true;
        }
        return false;
    }

    JsonValue value()
    {
        skip_ws();
        if (pos_ >= s_.size())
            error("unexpected end of input");
        const char c = s_[pos_];
        if (c == '{')
            return object();
        if (c == '[')
            return array();
        if (c == '"')
            return JsonValue::string(string());
        if (consume("true"))
            return JsonValue::boolean(true);
        if (consume("false"))
            return JsonValue::boolean(false);
        if (consume("null"))
            return JsonValue();
        return number();
    }

    JsonValue object()
    {
        JsonValue v = JsonValue::object();
        ++pos_;
        skip_ws();
        if (pos_ < s_.size() && s_[pos_] == '}')
        {
            ++pos_;
            return v;
        }
        for (;;)
        {
            skip_ws();
            if (pos_ >= s_.size() || s_[pos_] != '"')
                error("expected a key");
            std::string key = string();
            skip_ws();
            if (pos_ >= s_.size() || s_[pos_] != ':')
                error("expected ':'");
            ++pos_;
            v[key] = value();
            skip_ws();
            if (pos_ < s_.size() && s_[pos_] == ',')
            {
                ++pos_;
                continue;
            }
            if (pos_ < s_.size() && s_[pos_] == '}')
            {
                ++pos_;
                return v;
            }
            error("expected ',' or '}'");
        }
    }

    JsonValue array()
    {
        JsonValue v = JsonValue::array();
        ++pos_;
        skip_ws();
        if (pos_ < s_.size() && s_[pos_] == ']')
        {
            ++pos_;
            return v;
        }
        for (;;)
        {
            v.push_back(value());
            skip_ws();
            if (pos_ < s_.size() && s_[pos_] == ',')
            {
                ++pos_;
                continue;
            }
            if (pos_ < s_.size() && s_[pos_] == ']')
            {
                ++pos_;
                return v;
            }
            error("expected ',' or ']'");
        }
    }

    static void append_utf8(std::string &out, unsigned int cp)
    {
        if (cp < 0x80)
            out += static_cast<char>(cp);
        else if (cp < 0x800)
        {
            out += static_cast<char>(0xC0 | (cp >> 6));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
        else if (cp < 0x10000)
        {
            out += static_cast<char>(0xE0 | (cp >> 12));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
        else
        {
            out += static_cast<char>(0xF0 | (cp >> 18));
            out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }

    unsigned int hex4()
    {
        if (pos_ + 4 > s_.size())
            error("short \\u escape");
        unsigned int v = 0;
        for (int i = 0; i < 4; ++i)
        {
            const char c = s_[pos_++];
            v <<= 4;
            if (c >= '0' && c <= '9')
                v |= static_cast<unsigned int>(c - '0');
            else if (c >= 'a' && c <= 'f')
                v |= static_cast<unsigned int>(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F')
                v |= static_cast<unsigned int>(c - 'A' + 10);
            else
                error("bad \\u escape");
        }
        return v;
    }

    std::string string()
    {
        ++pos_;  // opening quote
        std::string out;
        while (pos_ < s_.size() && s_[pos_] != '"')
        {
            char c = s_[pos_++];
            if (c != '\\')
            {
                out += c;
                continue;
            }
            if (pos_ >= s_.size())
                error("unterminated escape");
            c = s_[pos_++];
            switch (c)
            {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u':
            {
                unsigned int cp = hex4();
                if (cp >= 0xD800 && cp < 0xDC00 && pos_ + 6 <= s_.size() && s_[pos_] == '\\' && s_[pos_ + 1] == 'u')
                {
                    pos_ += 2;
                    const unsigned int lo = hex4();
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                append_utf8(out, cp);
                break;
            }
            default: error("bad escape");
            }
        }
        if (pos_ >= s_.size())
            error("unterminated string");
        ++pos_;  // closing quote
        return out;
    }

    JsonValue number()
    {
        const std::size_t start = pos_;
        if (pos_ < s_.size() && s_[pos_] == '-')
            ++pos_;
        while (pos_ < s_.size() &&
               ((s_[pos_] >= '0' && s_[pos_] <= '9') || s_[pos_] == '.' || s_[pos_] == 'e' || s_[pos_] == 'E' ||
                s_[pos_] == '+' || s_[pos_] == '-'))
            ++pos_;
        const std::string text = s_.substr(start, pos_ - start);
        if (text.empty() || text == "-")
            error("expected a value");
        char *end = nullptr;
        const double d = std::strtod(text.c_str(), &end);
        if (!end || *end != '\0')
            error("malformed number \"" + text + "\"");
        JsonValue v = JsonValue::number(d);
        // keep the literal so that integers read back exactly
        if (text.find_first_of(".eE") == std::string::npos && text[0] != '-')
            v = JsonValue::unsigned_number(std::strtoull(text.c_str(), nullptr, 10));
        return v;
    }

    const std::string &s_;
    std::size_t pos_ = 0;
};
}  // namespace

JsonValue JsonValue::parse(const std::string &text) { return Parser(text).document(); }

// ---------------------------------------------------------------------------
// settings <-> Properties (file_interaction.cpp:11-69)
// ---------------------------------------------------------------------------
namespace
{
unsigned int read_uint(const JsonValue &v, const std::string &what)
{
    const std::uint64_t u = v.as_unsigned(what);
    if (u > std::numeric_limits<unsigned int>::max())
        throw UnknownPropertyException("JSON property \"" + what + "\" is out of range");
    return static_cast<unsigned int>(u);
}
}  // namespace

JsonValue properties_to_json_value(const core::Properties &p)
{
    JsonValue doc = JsonValue::object();
    JsonValue &alg = doc["algorithmic"];
    alg["workGroupSize"] = JsonValue::unsigned_number(p.work_group_size);
    alg["debugMode"] = JsonValue::boolean(p.debug_mode);
    alg["timeSteps"] = JsonValue::unsigned_number(p.time_steps);
    alg["dataLayout"] = JsonValue::string(p.data_layout);
    alg["algorithm"] = JsonValue::string(p.algorithm);
    alg["frameUpdateInterval"] = JsonValue::unsigned_number(p.frame_update_interval);
    JsonValue &dom = doc["domain"];
    dom["scenario"] = JsonValue::string(p.scenario);
    dom["verticalNodes"] = JsonValue::unsigned_number(p.vertical_nodes);
    dom["horizontalNodes"] = JsonValue::unsigned_number(p.horizontal_nodes);
    JsonValue &phy = doc["physical"];
    phy["inletVelocity"]["x"] = JsonValue::number(p.inlet_velocity_x);
    phy["inletVelocity"]["y"] = JsonValue::number(p.inlet_velocity_y);
    phy["inletDensity"] = JsonValue::number(p.inlet_density);
    phy["outletVelocity"]["x"] = JsonValue::number(p.outlet_velocity_x);
    phy["outletVelocity"]["y"] = JsonValue::number(p.outlet_velocity_y);
    phy["outletDensity"] = JsonValue::number(p.outlet_density);
    phy["relaxationTime"] = JsonValue::number(p.relaxation_time);
    if (!p.extensions.is_default())
    {
        const core::Extensions &e = p.extensions;
        JsonValue &x = doc["mi355x"];
        x["storage"] = JsonValue::string(e.storage);
        x["rounding"] = JsonValue::string(e.rounding);
        x["tileKernel"] = JsonValue::string(e.tile_kernel);
        x["nanProtection"] = JsonValue::boolean(e.nan_protection);
        x["seed"] = JsonValue::unsigned_number(e.seed);
        x["porosity"] = JsonValue::number(e.porosity);
        x["radiusMin"] = JsonValue::number(e.radius_min);
        x["radiusMax"] = JsonValue::number(e.radius_max);
        x["obstacleMargin"] = JsonValue::unsigned_number(e.obstacle_margin);
        x["maxAttempts"] = JsonValue::unsigned_number(e.max_attempts);
        x["cylinderX"] = JsonValue::unsigned_number(e.cylinder_x);
        x["cylinderRadius"] = JsonValue::number(e.cylinder_radius);
    }
    return doc;
}

void properties_to_json(const core::Properties &properties)
{
    std::ofstream file(properties.settings_path);
    if (!file)
        throw exceptions::Exception("cannot write settings file \"" + properties.settings_path + "\"");
    file << properties_to_json_value(properties).dump(4) << "\n";
}

core::Properties json_value_to_properties(const JsonValue &data, int offset)
{
    const JsonValue &alg = data.at("algorithmic");
    const JsonValue &dom = data.at("domain");
    const JsonValue &phy = data.at("physical");
    const long v = static_cast<long>(read_uint(dom.at("verticalNodes"), "domain.verticalNodes")) + offset;
    const long h = static_cast<long>(read_uint(dom.at("horizontalNodes"), "domain.horizontalNodes")) + offset;
    if (v < 0 || h < 0)
        throw exceptions::json::PropertyArgumentException("node counts must not be negative");
    core::Properties p(alg.at("algorithm").as_string("algorithmic.algorithm"),
                       alg.at("dataLayout").as_string("algorithmic.dataLayout"),
                       alg.at("debugMode").as_bool("algorithmic.debugMode"),
                       read_uint(alg.at("workGroupSize"), "algorithmic.workGroupSize"),
                       read_uint(alg.at("timeSteps"), "algorithmic.timeSteps"),
                       read_uint(alg.at("frameUpdateInterval"), "algorithmic.frameUpdateInterval"),
                       dom.at("scenario").as_string("domain.scenario"), static_cast<unsigned int>(v),
                       static_cast<unsigned int>(h), phy.at("inletVelocity").at("x").as_double("inletVelocity.x"),
                       phy.at("inletVelocity").at("y").as_double("inletVelocity.y"),
                       phy.at("inletDensity").as_double("physical.inletDensity"),
                       phy.at("outletVelocity").at("x").as_double("outletVelocity.x"),
                       phy.at("outletVelocity").at("y").as_double("outletVelocity.y"),
                       phy.at("outletDensity").as_double("physical.outletDensity"),
                       phy.at("relaxationTime").as_double("physical.relaxationTime"));
    if (const JsonValue *x = data.find("mi355x"))
    {
        core::Extensions &e = p.extensions;
        if (const JsonValue *f = x->find("storage"))
            e.storage = f->as_string("mi355x.storage");
        if (const JsonValue *f = x->find("rounding"))
            e.rounding = f->as_string("mi355x.rounding");
        if (const JsonValue *f = x->find("tileKernel"))
            e.tile_kernel = f->as_string("mi355x.tileKernel");
        if (const JsonValue *f = x->find("nanProtection"))
            e.nan_protection = f->as_bool("mi355x.nanProtection");
        if (const JsonValue *f = x->find("seed"))
            e.seed = f->as_unsigned("mi355x.seed");
        if (const JsonValue *f = x->find("porosity"))
            e.porosity = f->as_double("mi355x.porosity");
        if (const JsonValue *f = x->find("radiusMin"))
            e.radius_min = f->as_double("mi355x.radiusMin");
        if (const JsonValue *f = x->find("radiusMax"))
            e.radius_max = f->as_double("mi355x.radiusMax");
        if (const JsonValue *f = x->find("obstacleMargin"))
            e.obstacle_margin = read_uint(*f, "mi355x.obstacleMargin");
        if (const JsonValue *f = x->find("maxAttempts"))
            e.max_attempts = f->as_unsigned("mi355x.maxAttempts");
        if (const JsonValue *f = x->find("cylinderX"))
            e.cylinder_x = read_uint(*f, "mi355x.cylinderX");
        if (const JsonValue *f = x->find("cylinderRadius"))
            e.cylinder_radius = f->as_double("mi355x.cylinderRadius");
    }
    return p;
}

core::Properties json_to_properties(const std::string &path, int offset)
{
    std::ifstream file(path);
    if (!file)
        throw exceptions::json::MissingPropertyException("cannot open settings file \"" + path + "\"");
    std::stringstream buffer;
    buffer << file.rdbuf();
    core::Properties p = json_value_to_properties(JsonValue::parse(buffer.str()), offset);
    p.settings_path = path;
    return p;
}

}  // namespace file_interaction
}  // namespace lbm
