// OpenFOAM ASCII dictionaries and lists: the tokenizer, the dictionary tree and typed
// look-ups used by the case reader (constant/picProperties, system/controlDict,
// system/picInitialiseDict, system/blockMeshDict, 0/ fields, constant/polyMesh lists).
// Host code only; errors are thrown as FoamError and turned into status codes at the C-ABI.
#pragma once
#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace picmi {
namespace foam {

struct FoamError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Token {
    enum Kind { Word, Number, String, Punct };
    Kind kind = Word;
    std::string text;  // the word, the string's contents or the punctuation character
    double num = 0.0;
    bool integral = false;
    int line = 0;
    bool is(char c) const { return kind == Punct && text.size() == 1 && text[0] == c; }
};

// Splits OpenFOAM text into tokens: // and /* */ comments dropped; ( ) { } [ ] ; are
// punctuation; "..." strings; every other run of characters is a word, or a number when
// the whole run parses as one.
std::vector<Token> tokenize(const std::string& text, const std::string& where);

class Dict;

struct Entry {
    std::string key;
    std::vector<Token> value;    // a primitive entry's tokens, without the closing ';'
    std::shared_ptr<Dict> dict;  // a sub-dictionary, or null
};

// Sequential reader over a token range, for values and bare lists.
class Reader {
public:
    Reader(const std::vector<Token>& t, std::string where, size_t pos = 0, size_t end = (size_t)-1);
    bool done() const { return pos_ >= end_; }
    const Token& peek() const;
    const Token& next();
    void expect(char c);
    bool accept(char c);
    double number();
    long long label();
    std::string word();
    std::array<double, 3> vector3();
    // an optional element count before a list: "N (" or "(" ; returns -1 without a count
    long long list_open();
    size_t pos() const { return pos_; }
    [[noreturn]] void fail(const std::string& what) const;

private:
    const std::vector<Token>& t_;
    std::string where_;
    size_t pos_, end_;
};

class Dict {
public:
    std::string where;  // file (and sub-dictionary path) for messages
    std::vector<Entry> entries;

    const Entry* find(const std::string& key) const;
    bool has(const std::string& key) const { return find(key) != nullptr; }
    const Dict* sub(const std::string& key) const;      // null when absent or not a dictionary
    const Dict& sub_required(const std::string& key) const;
    // the entry's tokens as a Reader (throws when absent)
    Reader value(const std::string& key) const;

    double scalar(const std::string& key) const;
    double scalar_or(const std::string& key, double def) const;
    long long label(const std::string& key) const;
    std::string word(const std::string& key) const;
    std::string word_or(const std::string& key, const std::string& def) const;
    bool boolean(const std::string& key) const;
    bool boolean_or(const std::string& key, bool def) const;
    std::array<double, 3> vector3(const std::string& key) const;
    std::vector<std::string> words(const std::string& key) const;
    std::vector<double> scalars(const std::string& key) const;
    [[noreturn]] void fail(const std::string& key, const std::string& what) const;
};

// Parses "keyword value; keyword { ... }" entries from tokens[pos] up to the end, or up to
// the closing brace when braced (pos then points past it).
std::shared_ptr<Dict> parse_entries(const std::vector<Token>& t, size_t& pos, const std::string& where, bool braced);

// A whole dictionary file (FoamFile header included as an ordinary entry).
std::shared_ptr<Dict> read_dict_file(const std::string& path);

// A list file (constant/polyMesh/points, faces, owner, neighbour, boundary; lagrangian
// fields): the FoamFile header, then the tokens of the list that follows.
struct ListFile {
    std::shared_ptr<Dict> header;
    std::vector<Token> tokens;
    size_t body = 0;  // first token after the header
    std::string where;
};
ListFile read_list_file(const std::string& path);

bool file_exists(const std::string& path);
std::string read_text(const std::string& path);
// the FoamFile header block written in front of every file
std::string header(const std::string& cls, const std::string& location, const std::string& object);
// a double as OpenFOAM writes it at the given precision (general format)
std::string fmt(double v, int precision = 17);

}  // namespace foam
}  // namespace picmi
