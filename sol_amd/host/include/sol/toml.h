// A TOML 1.0 reader for scene files (the reference reads them with toml++, contrib/tomlplusplus).
// Supported: comments, bare / quoted / dotted keys, [tables], [[arrays of tables]], basic and
// literal strings (single- and multi-line, with escapes), integers (decimal, hex, octal, binary,
// '_' separators), floats (inf / nan included), booleans, arrays and inline tables. Date-time
// values are rejected. Errors throw sol::SourceError with the position.
#ifndef SOL_TOML_H
#define SOL_TOML_H

#include <cstdint>
#include <memory>
#include <optional>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

namespace sol::toml {

struct Value {
    enum class Kind { Table, Array, String, Integer, Float, Boolean };
    Kind kind = Kind::Table;
    std::string str;
    int64_t integer = 0;
    double real = 0.0;
    bool boolean = false;
    std::vector<std::unique_ptr<Value>> array;
    std::vector<std::pair<std::string, std::unique_ptr<Value>>> table;   // in file order
    size_t line = 0, col = 0;
    // Bookkeeping of the parser: tables named by a [header], arrays made by [[header]], inline
    // tables and arrays (closed to later additions).
    bool header_defined = false, table_array = false, sealed = false;

    explicit Value(Kind k) : kind(k) {}

    bool is_table() const { return kind == Kind::Table; }
    bool is_array() const { return kind == Kind::Array; }
    /// Member of a table, or nullptr.
    const Value* get(std::string_view key) const;
    /// Element of an array, or nullptr.
    const Value* at(size_t i) const { return is_array() && i < array.size() ? array[i].get() : nullptr; }
    /// Number as a float (integers convert), as toml++'s value_or<float> reads it.
    std::optional<double> number() const;
    std::optional<std::string> string() const;
};

/// Parses a whole document; the result is the root table.
std::unique_ptr<Value> parse(std::string_view text, const std::string& file_name);

} // namespace sol::toml

#endif
