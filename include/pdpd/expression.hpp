// expression.hpp -- parameter expressions of WorldState files, the language of
// the reference's include/dpd/core/expression.hpp: numbers, variables
// ([A-Za-z_][A-Za-z0-9_]*), + - * / with the usual precedence (left to right
// within a level) and parentheses. Numbers are read with strtod; '-' is always
// an operator, so a literal cannot start with one. Whitespace separates
// tokens. The text form puts every binary operation in parentheses and writes
// constants with std::to_string, so parsing and printing round-trips with the
// reference. Division by zero during evaluation throws. Space variables bound
// per cell by the engines: x, y, z (cell coordinates) and ux, uy, uz
// ((coordinate + 0.5) / cells along the axis).
#pragma once

#include <cctype>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace pdpd_expr {

struct Node {
    char op = 'c';                     // 'c' constant, 'v' variable, or one of + - * /
    double value = 0.0;
    std::string name;
    std::shared_ptr<const Node> lhs, rhs;

    double eval(const std::unordered_map<std::string, double> &b) const
    {
        switch (op) {
        case 'c': return value;
        case 'v': {
            auto it = b.find(name);
            if (it == b.end()) throw std::runtime_error("Could not find variable binding for '" + name + "'");
            return it->second;
        }
        case '+': return lhs->eval(b) + rhs->eval(b);
        case '-': return lhs->eval(b) - rhs->eval(b);
        case '*': return lhs->eval(b) * rhs->eval(b);
        default: {
            const double d = rhs->eval(b);
            if (d == 0) throw std::runtime_error("Divide by 0 in " + text());
            return lhs->eval(b) / d;
        }
        }
    }

    std::string text() const
    {
        if (op == 'c') return std::to_string(value);
        if (op == 'v') return name;
        return "(" + lhs->text() + std::string(1, op) + rhs->text() + ")";
    }

    void variables(std::unordered_set<std::string> &out) const
    {
        if (op == 'v') out.insert(name);
        if (lhs) lhs->variables(out);
        if (rhs) rhs->variables(out);
    }
};

using NodePtr = std::shared_ptr<const Node>;

class Parser {
public:
    explicit Parser(const std::string &src) { tokenize(src); }

    NodePtr parse()
    {
        if (m_tok.empty()) throw std::runtime_error("empty expression");
        NodePtr e = sum();
        if (m_pos != m_tok.size()) throw std::runtime_error("Unconsumed tokens.");
        return e;
    }

private:
    struct Tok {
        bool number;
        double value;
        std::string text;
    };

    void tokenize(const std::string &src)
    {
        const char *p = src.c_str();
        while (*p) {
            const char c = *p;
            if (std::isspace((unsigned char)c)) {
                p++;
            } else if (c == '(' || c == ')' || c == '+' || c == '-' || c == '*' || c == '/') {
                m_tok.push_back({false, 0.0, std::string(1, c)});
                p++;
            } else if (std::isalpha((unsigned char)c) || c == '_') {
                const char *e = p + 1;
                while (std::isalnum((unsigned char)*e) || *e == '_') e++;
                m_tok.push_back({false, 0.0, std::string(p, e)});
                p = e;
            } else {
                char *e = nullptr;
                const double v = std::strtod(p, &e);
                if (e == p) throw std::runtime_error("Couldn't parse expression from " + std::string(p));
                m_tok.push_back({true, v, std::string(p, (const char *)e)});
                p = e;
            }
        }
    }

    bool at(const char *s) const { return m_pos < m_tok.size() && !m_tok[m_pos].number && m_tok[m_pos].text == s; }

    static NodePtr binary(char op, NodePtr l, NodePtr r)
    {
        auto n = std::make_shared<Node>();
        n->op = op;
        n->lhs = std::move(l);
        n->rhs = std::move(r);
        return n;
    }

    NodePtr sum()
    {
        NodePtr e = product();
        while (m_pos < m_tok.size() && !at(")")) {
            if (at("+") || at("-")) {
                const char op = m_tok[m_pos++].text[0];
                e = binary(op, e, product());
            } else {
                throw std::runtime_error("Unexpected token in parse_expression_terms.");
            }
        }
        return e;
    }

    NodePtr product()
    {
        NodePtr e = atom();
        while (m_pos < m_tok.size() && !at(")") && !at("+") && !at("-")) {
            if (at("*") || at("/")) {
                const char op = m_tok[m_pos++].text[0];
                e = binary(op, e, atom());
            } else {
                throw std::runtime_error("Unexpected token in parse_expression_factors.");
            }
        }
        return e;
    }

    NodePtr atom()
    {
        if (m_pos >= m_tok.size()) throw std::runtime_error("expression ends early");
        const Tok &t = m_tok[m_pos++];
        auto n = std::make_shared<Node>();
        if (t.number) {
            n->value = t.value;
            return n;
        }
        if (std::isalpha((unsigned char)t.text[0]) || t.text[0] == '_') {
            n->op = 'v';
            n->name = t.text;
            return n;
        }
        if (t.text == "(") {
            NodePtr e = sum();
            if (!at(")")) throw std::runtime_error("Un-matched '(' bracket.");
            m_pos++;
            return e;
        }
        throw std::runtime_error("Unexpected token in parse_expression_base, token = " + t.text);
    }

    std::vector<Tok> m_tok;
    size_t m_pos = 0;
};

inline NodePtr parse(const std::string &src) { return Parser(src).parse(); }

}  // namespace pdpd_expr
