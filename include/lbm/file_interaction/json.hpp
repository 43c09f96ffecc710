// Minimal JSON document model for the settings files: null, bool, number,
// string, array, object (insertion-ordered).  Parsing follows RFC 8259;
// numbers are kept as double plus the literal text so that unsigned values
// read back exactly.  Serialisation indents by a given width.
#ifndef LBM_JSON_HPP
#define LBM_JSON_HPP

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace lbm
{
namespace file_interaction
{

class JsonValue
{
  public:
    enum class Kind
    {
        Null,
        Bool,
        Number,
        String,
        Array,
        Object
    };

    JsonValue() = default;
    static JsonValue boolean(bool b);
    static JsonValue number(double d);
    static JsonValue unsigned_number(std::uint64_t u);
    static JsonValue string(std::string s);
    static JsonValue array();
    static JsonValue object();

    Kind kind() const { return kind_; }
    bool is_object() const { return kind_ == Kind::Object; }

    // Typed reads; throw exceptions::json::UnknownPropertyException on a
    // kind mismatch, naming `what`.
    bool as_bool(const std::string &what) const;
    double as_double(const std::string &what) const;
    std::uint64_t as_unsigned(const std::string &what) const;
    const std::string &as_string(const std::string &what) const;

    // Object access.  at() throws exceptions::json::MissingPropertyException.
    const JsonValue &at(const std::string &key) const;
    const JsonValue *find(const std::string &key) const;
    JsonValue &operator[](const std::string &key);  // inserts null if absent

    // Array access.
    void push_back(JsonValue v);
    const std::vector<JsonValue> &items() const { return array_; }
    JsonValue &item(std::size_t i) { return array_.at(i); }

    std::string dump(int indent = 4) const;
    static JsonValue parse(const std::string &text);  // throws UnknownPropertyException on syntax errors

  private:
    void dump_to(std::string &out, int indent, int level) const;

    Kind kind_ = Kind::Null;
    bool bool_ = false;
    double number_ = 0.0;
    std::string text_;  // string value, or the literal of a number
    std::vector<JsonValue> array_;
    std::vector<std::pair<std::string, JsonValue>> object_;
};

}  // namespace file_interaction
}  // namespace lbm

#endif  // LBM_JSON_HPP
