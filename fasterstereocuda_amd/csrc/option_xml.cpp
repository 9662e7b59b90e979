// Option files of the stereo console (see option_xml.h): a small
// recursive-descent XML reader (elements, character data, the five
// predefined entities and numeric character references, attributes skipped,
// comments / processing instructions / CDATA / DOCTYPE handled) and the
// writer of the two option forms.
#include "option_xml.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <sys/stat.h>

namespace fsc {

const XmlElement* XmlElement::child(const char* n) const {
    for (const XmlElement& c : children)
        if (c.name == n) return &c;
    return nullptr;
}

namespace {

class Parser {
public:
    Parser(const std::string& s, std::string* err) : s_(s), err_(err) {}

    bool document(XmlElement* root) {
        if (!misc()) return false;
        if (!at("<")) return fail("no root element");
        if (!element(root, 0)) return false;
        if (!misc()) return false;
        if (p_ != s_.size()) return fail("text after the root element");
        return true;
    }

private:
    const std::string& s_;
    std::string* err_;
    size_t p_ = 0;

    bool fail(const std::string& m) {
        if (err_) *err_ = m + " (offset " + std::to_string(p_) + ")";
        return false;
    }
    bool at(const char* t) const { return s_.compare(p_, std::char_traits<char>::length(t), t) == 0; }
    static bool blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }
    static bool name_char(char c) {
        return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') ||
               c == '_' || c == '-' || c == '.' || c == ':' || (unsigned char)c >= 0x80;
    }
    void skip_blanks() {
        while (p_ < s_.size() && blank(s_[p_])) ++p_;
    }
    bool skip_past(const char* end) {
        const size_t q = s_.find(end, p_);
        if (q == std::string::npos) return fail(std::string("unterminated construct, no ") + end);
        p_ = q + std::char_traits<char>::length(end);
        return true;
    }
    // blanks, comments, processing instructions and a DOCTYPE outside the root
    bool misc() {
        while (true) {
            skip_blanks();
            if (at("<?")) {
                if (!skip_past("?>")) return false;
            } else if (at("<!--")) {
                if (!skip_past("-->")) return false;
            } else if (at("<!DOCTYPE")) {
                if (!skip_past(">")) return false;
            } else {
                return true;
            }
        }
    }
    bool name(std::string* out) {
        const size_t b = p_;
        while (p_ < s_.size() && name_char(s_[p_])) ++p_;
        if (p_ == b) return fail("expected a name");
        *out = s_.substr(b, p_ - b);
        return true;
    }
    static void put_utf8(unsigned long cp, std::string* out) {
        if (cp < 0x80) {
            out->push_back((char)cp);
        } else if (cp < 0x800) {
            out->push_back((char)(0xC0 | (cp >> 6)));
            out->push_back((char)(0x80 | (cp & 0x3F)));
        } else if (cp < 0x10000) {
            out->push_back((char)(0xE0 | (cp >> 12)));
            out->push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
            out->push_back((char)(0x80 | (cp & 0x3F)));
        } else {
            out->push_back((char)(0xF0 | (cp >> 18)));
            out->push_back((char)(0x80 | ((cp >> 12) & 0x3F)));
            out->push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
            out->push_back((char)(0x80 | (cp & 0x3F)));
        }
    }
    bool entity(std::string* out) {
        const size_t semi = s_.find(';', p_);
        if (semi == std::string::npos || semi - p_ > 12) return fail("bad entity");
        const std::string e = s_.substr(p_ + 1, semi - p_ - 1);
        if (e == "lt") out->push_back('<');
        else if (e == "gt") out->push_back('>');
        else if (e == "amp") out->push_back('&');
        else if (e == "quot") out->push_back('"');
        else if (e == "apos") out->push_back('\'');
        else if (e.size() > 1 && e[0] == '#') {
            const bool hex = e[1] == 'x' || e[1] == 'X';
            char* end = nullptr;
            const char* digits = e.c_str() + (hex ? 2 : 1);
            const unsigned long cp = std::strtoul(digits, &end, hex ? 16 : 10);
            if (end == digits || *end != '\0' || cp > 0x10FFFF) return fail("bad character reference");
            put_utf8(cp, out);
        } else {
            return fail("unknown entity &" + e + ";");
        }
        p_ = semi + 1;
        return true;
    }
    bool attributes() {
        while (true) {
            skip_blanks();
            if (p_ >= s_.size()) return fail("unterminated start tag");
            if (s_[p_] == '>' || at("/>")) return true;
            std::string n;
            if (!name(&n)) return false;
            skip_blanks();
            if (p_ >= s_.size() || s_[p_] != '=') return fail("attribute without a value");
            ++p_;
            skip_blanks();
            if (p_ >= s_.size() || (s_[p_] != '"' && s_[p_] != '\'')) return fail("unquoted attribute");
            const char q = s_[p_++];
            const size_t e = s_.find(q, p_);
            if (e == std::string::npos) return fail("unterminated attribute value");
            p_ = e + 1;
        }
    }
    bool element(XmlElement* el, int depth) {
        if (depth > 256) return fail("elements nested too deeply");
        ++p_;   // '<'
        if (!name(&el->name)) return false;
        if (!attributes()) return false;
        if (at("/>")) {
            p_ += 2;
            return true;
        }
        ++p_;   // '>'
        while (true) {
            if (p_ >= s_.size()) return fail("element <" + el->name + "> is not closed");
            const char c = s_[p_];
            if (c == '&') {
                if (!entity(&el->text)) return false;
            } else if (c != '<') {
                el->text.push_back(c);
                ++p_;
            } else if (at("</")) {
                p_ += 2;
                std::string n;
                if (!name(&n)) return false;
                if (n != el->name) return fail("</" + n + "> closes <" + el->name + ">");
                skip_blanks();
                if (p_ >= s_.size() || s_[p_] != '>') return fail("bad end tag");
                ++p_;
                return true;
            } else if (at("<!--")) {
                if (!skip_past("-->")) return false;
            } else if (at("<![CDATA[")) {
                const size_t b = p_ + 9;
                if (!skip_past("]]>")) return false;
                el->text.append(s_, b, p_ - 3 - b);
            } else if (at("<?")) {
                if (!skip_past("?>")) return false;
            } else {
                el->children.emplace_back();
                if (!element(&el->children.back(), depth + 1)) return false;
            }
        }
    }
};

bool exists(const std::string& path) {
    struct stat st;
    return ::stat(path.c_str(), &st) == 0;
}

bool read_file(const std::string& path, std::string* out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    *out = ss.str();
    return true;
}

bool load_root(const std::string& path, XmlElement* root) {
    std::string doc, err;
    if (!read_file(path, &doc)) return false;
    if (!parse_xml(doc, root, &err)) {
        std::fprintf(stderr, "%s: %s\n", path.c_str(), err.c_str());
        return false;
    }
    return true;
}

// atoi / atof of an element's text, as the reference reads every value
template <class T>
void read_int(const XmlElement* sec, const char* key, T* v) {
    if (const XmlElement* e = sec ? sec->child(key) : nullptr) *v = (T)std::atoi(e->text.c_str());
}
void read_bool(const XmlElement* sec, const char* key, bool* v) {
    if (const XmlElement* e = sec ? sec->child(key) : nullptr) *v = std::atoi(e->text.c_str()) != 0;
}
template <class T>
void read_real(const XmlElement* sec, const char* key, T* v) {
    if (const XmlElement* e = sec ? sec->child(key) : nullptr) *v = (T)std::atof(e->text.c_str());
}

class Writer {
public:
    explicit Writer(const char* root) : root_(root) {
        out_ << "<?xml version=\"1.0\" encoding=\"UTF-8\" standalone=\"no\"?>\n<" << root << ">\n";
    }
    void open(const char* sec) { out_ << "    <" << sec << ">\n"; }
    void close(const char* sec) { out_ << "    </" << sec << ">\n"; }
    void field(const char* key, const std::string& v) {
        out_ << "        <" << key << ">" << v << "</" << key << ">\n";
    }
    bool save(const std::string& path) {
        out_ << "</" << root_ << ">\n";
        std::ofstream f(path, std::ios::binary);
        if (!f) return false;
        f << out_.str();
        return bool(f);
    }

private:
    const char* root_;
    std::ostringstream out_;
};

template <class Opt>
void write_common(Writer& w, const Opt& o) {
    w.field("layers_num", std::to_string(o.num_layers));
    w.field("do_consistency_check", std::to_string(o.do_lr_check));
    w.field("do_remove_peaks", std::to_string(o.do_rm_peaks));
    w.field("do_smooth", std::to_string(o.do_smooth));
}

template <class Opt>
void read_common(const XmlElement& root, Opt& o) {
    const XmlElement* view = root.child("View");
    read_int(view, "width", &o.width);
    read_int(view, "height", &o.height);
    const XmlElement* ste = root.child("Stereo");
    read_int(ste, "layers_num", &o.num_layers);
    read_bool(ste, "do_consistency_check", &o.do_lr_check);
    read_bool(ste, "do_remove_peaks", &o.do_rm_peaks);
    read_bool(ste, "do_smooth", &o.do_smooth);
}

}  // namespace

bool parse_xml(const std::string& doc, XmlElement* root, std::string* err) {
    *root = XmlElement{};
    Parser p(doc, err);
    return p.document(root);
}

OptionKind option_kind(const std::string& path) {
    XmlElement root;
    if (!exists(path) || !load_root(path, &root)) return OptionKind::kNone;
    if (root.name == "Option1") return OptionKind::kDisparity;
    if (root.name == "Option2") return OptionKind::kDepth;
    return OptionKind::kNone;
}

bool load_option(const std::string& path, FasterStereoCuda::ste_opt1_t& o) {
    if (!exists(path)) {
        if (save_option(path, o)) std::printf("option file was created in %s\n", path.c_str());
        return false;
    }
    XmlElement root;
    if (!load_root(path, &root)) return false;
    read_common(root, o);
    const XmlElement* ste = root.child("Stereo");
    read_int(ste, "disp_min", &o.min_disp);
    read_int(ste, "disp_max", &o.max_disp);
    return true;
}

bool load_option(const std::string& path, FasterStereoCuda::ste_opt2_t& o) {
    if (!exists(path)) {
        if (save_option(path, o)) std::printf("option file was created in %s\n", path.c_str());
        return false;
    }
    XmlElement root;
    if (!load_root(path, &root)) return false;
    read_common(root, o);
    const XmlElement* ste = root.child("Stereo");
    read_real(ste, "depth_min", &o.min_depth);
    read_real(ste, "depth_max", &o.max_depth);
    const XmlElement* cam = root.child("Camera");
    read_real(cam, "x0_left_pix", &o.epi.x0_left);
    read_real(cam, "y0_left_pix", &o.epi.y0_left);
    read_real(cam, "x0_right_pix", &o.epi.x0_right);
    read_real(cam, "y0_right_pix", &o.epi.y0_right);
    read_real(cam, "baseline_mm", &o.epi.baseline);
    read_real(cam, "focus_pix", &o.epi.focus);
    return true;
}

bool save_option(const std::string& path, const FasterStereoCuda::ste_opt1_t& o) {
    Writer w("Option1");
    w.open("View");
    w.field("width", std::to_string(o.width));
    w.field("height", std::to_string(o.height));
    w.close("View");
    w.open("Stereo");
    w.field("disp_min", std::to_string(o.min_disp));
    w.field("disp_max", std::to_string(o.max_disp));
    write_common(w, o);
    w.close("Stereo");
    return w.save(path);
}

bool save_option(const std::string& path, const FasterStereoCuda::ste_opt2_t& o) {
    Writer w("Option2");
    w.open("View");
    w.field("width", std::to_string(o.width));
    w.field("height", std::to_string(o.height));
    w.close("View");
    w.open("Stereo");
    w.field("depth_min", std::to_string(o.min_depth));
    w.field("depth_max", std::to_string(o.max_depth));
    write_common(w, o);
    w.close("Stereo");
    w.open("Camera");
    w.field("x0_left_pix", std::to_string(o.epi.x0_left));
    w.field("y0_left_pix", std::to_string(o.epi.y0_left));
    w.field("x0_right_pix", std::to_string(o.epi.x0_right));
    w.field("y0_right_pix", std::to_string(o.epi.y0_right));
    w.field("baseline_mm", std::to_string(o.epi.baseline));
    w.field("focus_pix", std::to_string(o.epi.focus));
    w.close("Camera");
    return w.save(path);
}

}  // namespace fsc
