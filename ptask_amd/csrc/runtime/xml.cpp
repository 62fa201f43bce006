// xml.cpp -- Graph::Serialize / Graph::Deserialize: a graph's structure as an XML document
// (reference roles: XMLWriter.cpp:124-620, XMLReader.cpp, the Recorder's actions Recorder.cpp).
//
// Document layout (element names follow the reference's writer, so a file it wrote reads here):
//   <Graph>
//     <DatablockTemplates><Count/> <DatablockTemplate>Name Stride ParamType ElementsX/Y/Z
//         IsRecordStream IsByteAddressable [Pitch] [BufferDimensions*] [InitialValue]
//     <CompiledKernels><Count/> <CompiledKernel>KernelID SourceFile Operation
//     <Tasks><Count/> <Task>Name KernelID AcceleratorClass InputPorts OutputPorts [Priority]
//         Port: UniqueID OriginalIndex Type DatablockTemplate VariableBinding ParameterIndex
//               InOutIndex [MetaFunction PermanentBlock Marshallable]
//     <Channels><Count/> <Channel>Name Type SourcePort DestinationPort [Switch Capacity
//         ViewMaterialization InitialControlSignal Members]
//     <Actions><Count/> <Action>ActionName ... : BindControlPort, BindControlPropagationPort,
//         BindControlPropagationChannel, BindDescriptorPort, BindIterationScope,
//         SetComputeGeometry, SetBlockAndGridSize, SetPredicationType
// Elements in [] are written here and optional on reading (the reference's writer has no
// counterpart for them); ports are referred to by UniqueID and channels by Name.  The reference
// records the binding calls as they are made; here they are read off the graph's own state, in
// a fixed order, so serializing a deserialized graph writes the same text again.

#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "ptk_internal.h"

namespace PTask {

namespace {

struct XmlError : std::runtime_error {
    explicit XmlError(const std::string &s) : std::runtime_error(s) {}
};

// =============================================================================================
// Writing
// =============================================================================================

std::string Escape(const std::string &s) {
    std::string r;
    for (char c : s) {
        switch (c) {
            case '<': r += "&lt;"; break;
            case '>': r += "&gt;"; break;
            case '&': r += "&amp;"; break;
            case '"': r += "&quot;"; break;
            case '\'': r += "&apos;"; break;
            default: r += c;
        }
    }
    return r;
}

class XmlOut {
public:
    void Open(const char *tag) {
        Indent();
        m_s += "<";
        m_s += tag;
        m_s += ">\n";
        m_open.push_back(tag);
    }
    void Close() {
        std::string tag = m_open.back();
        m_open.pop_back();
        Indent();
        m_s += "</" + tag + ">\n";
    }
    void Text(const char *tag, const std::string &v) {
        Indent();
        m_s += std::string("<") + tag + ">" + Escape(v) + "</" + tag + ">\n";
    }
    void Int(const char *tag, long long v) { Text(tag, std::to_string(v)); }
    void UInt(const char *tag, unsigned long long v) { Text(tag, std::to_string(v)); }
    void Bool(const char *tag, bool v) { Text(tag, v ? "true" : "false"); }
    void Comment(const char *text) {
        Indent();
        m_s += std::string("<!--") + text + "-->\n";
    }
    const std::string &Str() const { return m_s; }

private:
    void Indent() { m_s.append(2 * m_open.size(), ' '); }
    std::string m_s = "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n";
    std::vector<std::string> m_open;
};

std::string Hex(const void *p, size_t n) {
    static const char d[] = "0123456789abcdef";
    std::string r;
    r.reserve(2 * n);
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; ++i) {
        r += d[b[i] >> 4];
        r += d[b[i] & 15];
    }
    return r;
}

bool IsScalarType(PTASK_PARM_TYPE t) { return t != PTPARM_BUFFER; }

std::string Fmt(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
std::string Fmt(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return buf;
}

// =============================================================================================
// Reading: a small XML parser (elements, text, comments, declarations, CDATA, the five named
// entities and numeric character references; attributes are skipped) into a tree.
// =============================================================================================

struct XmlNode {
    std::string tag;
    std::string text;
    std::vector<XmlNode> kids;

    const XmlNode *Find(const char *t) const {
        for (const XmlNode &k : kids)
            if (k.tag == t) return &k;
        return nullptr;
    }
    const XmlNode &Need(const char *t) const {
        const XmlNode *k = Find(t);
        if (!k) throw XmlError("<" + tag + "> has no <" + t + ">");
        return *k;
    }
    std::vector<const XmlNode *> All(const char *t) const {
        std::vector<const XmlNode *> r;
        for (const XmlNode &k : kids)
            if (k.tag == t) r.push_back(&k);
        return r;
    }
    std::string Str(const char *t) const { return Need(t).text; }
    long long Int(const char *t) const {
        const std::string &v = Need(t).text;
        errno = 0;
        char *end = nullptr;
        long long x = strtoll(v.c_str(), &end, 10);
        if (v.empty() || *end || errno) throw XmlError("<" + std::string(t) + "> in <" + tag + "> is not an integer: '" + v + "'");
        return x;
    }
    unsigned long long UInt(const char *t) const {
        const std::string &v = Need(t).text;
        if (!v.empty() && v[0] == '-') return (unsigned long long)Int(t);  // -1 for "none"
        errno = 0;
        char *end = nullptr;
        unsigned long long x = strtoull(v.c_str(), &end, 10);
        if (v.empty() || *end || errno) throw XmlError("<" + std::string(t) + "> in <" + tag + "> is not an integer: '" + v + "'");
        return x;
    }
    UINT Uint32(const char *t) const {
        unsigned long long x = UInt(t);
        if (x > 0xFFFFFFFFull) throw XmlError("<" + std::string(t) + "> in <" + tag + "> is out of range");
        return (UINT)x;
    }
    bool Bool(const char *t) const {
        const std::string &v = Need(t).text;
        if (v == "true" || v == "1") return true;
        if (v == "false" || v == "0") return false;
        throw XmlError("<" + std::string(t) + "> in <" + tag + "> is not a boolean: '" + v + "'");
    }
    // Optional elements written here only; `dflt` when absent.
    long long IntOr(const char *t, long long dflt) const { return Find(t) ? Int(t) : dflt; }
    unsigned long long UIntOr(const char *t, unsigned long long dflt) const { return Find(t) ? UInt(t) : dflt; }
    bool BoolOr(const char *t, bool dflt) const { return Find(t) ? Bool(t) : dflt; }
    // The children named `item` of a list element whose <Count> must agree with them.
    std::vector<const XmlNode *> List(const char *item) const {
        std::vector<const XmlNode *> r = All(item);
        if (Find("Count") && Int("Count") != (long long)r.size())
            throw XmlError("<" + tag + "> has Count " + Str("Count") + " but " + std::to_string(r.size()) + " <" + item + ">");
        return r;
    }
};

class XmlParser {
public:
    explicit XmlParser(const std::string &s) : m_s(s), m_i(0) {}
    // A graph document nests at most a handful of levels; deeper input is refused instead of
    // recursing without bound.
    static constexpr int kMaxDepth = 256;

    void Document(XmlNode *root) {
        Misc();
        if (!At("<")) Fail("expected the root element");
        Element(root);
        Misc();
        if (m_i != m_s.size()) Fail("text after the root element");
    }

private:
    [[noreturn]] void Fail(const std::string &what) {
        size_t line = 1;
        for (size_t k = 0; k < m_i && k < m_s.size(); ++k) line += m_s[k] == '\n';
        throw XmlError("XML line " + std::to_string(line) + ": " + what);
    }
    bool At(const char *p) const { return m_s.compare(m_i, strlen(p), p) == 0; }
    void Skip(const char *close) {
        size_t e = m_s.find(close, m_i);
        if (e == std::string::npos) Fail(std::string("unterminated construct, expected ") + close);
        m_i = e + strlen(close);
    }
    void Ws() {
        while (m_i < m_s.size() && strchr(" \t\r\n", m_s[m_i])) ++m_i;
    }
    // whitespace, comments, declarations and processing instructions between elements
    void Misc() {
        for (;;) {
            Ws();
            if (At("<?")) Skip("?>");
            else if (At("<!--")) Skip("-->");
            else if (At("<!DOCTYPE")) Skip(">");
            else return;
        }
    }
    std::string Name() {
        size_t b = m_i;
        while (m_i < m_s.size() && !strchr(" \t\r\n/>", m_s[m_i])) ++m_i;
        if (m_i == b) Fail("expected a name");
        return m_s.substr(b, m_i - b);
    }
    void Entity(std::string *out) {
        size_t e = m_s.find(';', m_i);
        if (e == std::string::npos || e - m_i > 12) Fail("bad character reference");
        std::string ent = m_s.substr(m_i + 1, e - m_i - 1);
        m_i = e + 1;
        if (ent == "lt") *out += '<';
        else if (ent == "gt") *out += '>';
        else if (ent == "amp") *out += '&';
        else if (ent == "quot") *out += '"';
        else if (ent == "apos") *out += '\'';
        else if (!ent.empty() && ent[0] == '#') {
            const bool hex = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X');
            const char *digits = ent.c_str() + (hex ? 2 : 1);
            char *end = nullptr;
            const unsigned long c = strtoul(digits, &end, hex ? 16 : 10);
            // a character reference names one Unicode scalar value: digits only, not NUL, not a
            // surrogate, at most U+10FFFF
            if (!*digits || *end || c == 0 || c > 0x10FFFF || (c >= 0xD800 && c <= 0xDFFF))
                Fail("bad character reference &" + ent + ";");
            if (c < 0x80) *out += (char)c;
            else if (c < 0x800) { *out += (char)(0xC0 | (c >> 6)); *out += (char)(0x80 | (c & 63)); }
            else if (c < 0x10000) { *out += (char)(0xE0 | (c >> 12)); *out += (char)(0x80 | ((c >> 6) & 63)); *out += (char)(0x80 | (c & 63)); }
            else { *out += (char)(0xF0 | (c >> 18)); *out += (char)(0x80 | ((c >> 12) & 63)); *out += (char)(0x80 | ((c >> 6) & 63)); *out += (char)(0x80 | (c & 63)); }
        } else {
            Fail("unknown entity &" + ent + ";");
        }
    }
    void Element(XmlNode *n, int depth = 0) {
        if (depth >= kMaxDepth) Fail("elements nested deeper than " + std::to_string(kMaxDepth) + " levels");
        ++m_i;  // '<'
        n->tag = Name();
        // attributes: skipped, quoted values may hold '>'
        while (m_i < m_s.size() && m_s[m_i] != '>' && !At("/>")) {
            char c = m_s[m_i++];
            if (c == '"' || c == '\'') {
                size_t e = m_s.find(c, m_i);
                if (e == std::string::npos) Fail("unterminated attribute value");
                m_i = e + 1;
            }
        }
        if (m_i >= m_s.size()) Fail("unterminated <" + n->tag + ">");
        if (At("/>")) {
            m_i += 2;
            return;
        }
        ++m_i;  // '>'
        std::string text;
        for (;;) {
            if (m_i >= m_s.size()) Fail("<" + n->tag + "> is not closed");
            if (At("</")) {
                m_i += 2;
                std::string close = Name();
                if (close != n->tag) Fail("</" + close + "> closes <" + n->tag + ">");
                Ws();
                if (!At(">")) Fail("expected '>'");
                ++m_i;
                break;
            }
            if (At("<!--")) Skip("-->");
            else if (At("<![CDATA[")) {
                size_t b = m_i + 9, e = m_s.find("]]>", b);
                if (e == std::string::npos) Fail("unterminated CDATA");
                text.append(m_s, b, e - b);
                m_i = e + 3;
            } else if (At("<?")) Skip("?>");
            else if (At("<")) {
                n->kids.emplace_back();
                Element(&n->kids.back(), depth + 1);
            } else if (m_s[m_i] == '&') Entity(&text);
            else text += m_s[m_i++];
        }
        // leaf text without the surrounding layout whitespace
        size_t b = text.find_first_not_of(" \t\r\n"), e = text.find_last_not_of(" \t\r\n");
        n->text = b == std::string::npos ? std::string() : text.substr(b, e - b + 1);
    }

    const std::string &m_s;
    size_t m_i;
};

std::vector<unsigned char> Unhex(const std::string &s) {
    if (s.size() % 2) throw XmlError("InitialValue bytes have odd length");
    std::vector<unsigned char> r(s.size() / 2);
    for (size_t i = 0; i < r.size(); ++i) {
        auto nib = [&](char c) -> int {
            if (c >= '0' && c <= '9') return c - '0';
            if (c >= 'a' && c <= 'f') return c - 'a' + 10;
            if (c >= 'A' && c <= 'F') return c - 'A' + 10;
            throw XmlError("InitialValue bytes are not hexadecimal");
        };
        r[i] = (unsigned char)(nib(s[2 * i]) << 4 | nib(s[2 * i + 1]));
    }
    return r;
}

// Everything of a template that Serialize writes, for comparing a file's template with one the
// runtime already holds under the same name.
std::string TemplateSignature(DatablockTemplate *t) {
    std::string s = std::to_string((int)t->GetParameterBaseType()) + "|" + std::to_string(t->DescribesRecordStream() ? 1 : 0) +
                    std::to_string(t->IsByteAddressable() ? 1 : 0);
    for (UINT c = 0; c < NUM_DATABLOCK_CHANNELS; ++c) {
        BUFFERDIMENSIONS d = t->GetBufferDimensions(c);
        s += "|" + std::to_string(d.uiXElements) + "," + std::to_string(d.uiYElements) + "," + std::to_string(d.uiZElements) +
             "," + std::to_string(d.cbElementStride) + "," + std::to_string(d.cbPitch);
    }
    if (t->HasInitialValue())
        s += "|init" + std::to_string(t->GetInitialValueElements()) + (t->IsInitialValueEmpty() ? "e" : "") +
             Hex(t->GetInitialValue(), t->GetInitialValueSizeBytes());
    return s;
}

}  // namespace

// =============================================================================================
// Graph::Serialize
// =============================================================================================

// The document for this graph (graph lock held); throws XmlError for what has no text form.
std::string Graph::SerializeLocked() {

    // every port of every task, inputs then outputs; UIDs name ports in the document
    std::vector<Port *> ports;
    std::map<UINT, Port *> byUid;
    for (Task *t : m_vTasks) {
        for (UINT i = 0; i < t->GetInputPortCount(); ++i) ports.push_back(t->GetInputPort(i));
        for (UINT i = 0; i < t->GetOutputPortCount(); ++i) ports.push_back(t->GetOutputPort(i));
    }
    for (Port *p : ports) {
        auto it = byUid.find(p->GetUID());
        if (it != byUid.end() && it->second != p)
            throw XmlError(Fmt("ports %s and %s share UniqueID %u", it->second->GetVariableBinding(),
                                        p->GetVariableBinding(), p->GetUID()));
        byUid[p->GetUID()] = p;
    }
    auto known = [&](Port *p) { return p && byUid.count(p->GetUID()) && byUid[p->GetUID()] == p; };

    // templates and kernels in order of first use, so the text is deterministic
    std::vector<DatablockTemplate *> templates;
    std::set<DatablockTemplate *> seenT;
    for (Port *p : ports)
        if (p->GetTemplate() && seenT.insert(p->GetTemplate()).second) templates.push_back(p->GetTemplate());
    for (Channel *c : m_vChannels)
        if (c->GetTemplate() && seenT.insert(c->GetTemplate()).second) templates.push_back(c->GetTemplate());
    std::vector<CompiledKernel *> kernels;
    std::map<CompiledKernel *, int> kernelId;
    for (Task *t : m_vTasks)
        if (!kernelId.count(t->GetCompiledKernel())) {
            kernelId[t->GetCompiledKernel()] = (int)kernels.size();
            kernels.push_back(t->GetCompiledKernel());
        }
    std::set<std::string> names;
    for (Channel *c : m_vChannels)
        if (!names.insert(c->GetName()).second)
            throw XmlError(Fmt("two channels are named '%s'", c->GetName()));
    std::set<std::string> taskNames;
    for (Task *t : m_vTasks)
        if (!taskNames.insert(t->GetTaskName()).second)
            throw XmlError(Fmt("two tasks are named '%s'", t->GetTaskName()));
    std::set<Channel *> members;
    for (Channel *c : m_vChannels)
        if (c->GetType() == CT_MULTI)
            for (Channel *m : static_cast<MultiChannel *>(c)->GetMembers()) members.insert(m);

    XmlOut w;
    w.Open("Graph");
    w.Text("Name", m_strName);

    w.Comment("  DatablockTemplates  ");
    w.Open("DatablockTemplates");
    w.UInt("Count", templates.size());
    for (DatablockTemplate *t : templates) {
        BUFFERDIMENSIONS d = t->GetBufferDimensions(DBDATA_IDX);
        if (IsScalarType(t->GetParameterBaseType())) {
            bool plain = d.uiXElements == 1 && d.uiYElements == 1 && d.uiZElements == 1 && d.cbPitch == d.cbElementStride &&
                         !t->DescribesRecordStream() && !t->IsByteAddressable();
            for (UINT c = 1; c < NUM_DATABLOCK_CHANNELS; ++c) plain = plain && t->GetBufferDimensions(c).AllocationSizeBytes() == 0;
            if (!plain)
                throw XmlError(Fmt("template %s: a parameter template with buffer dimensions has no text form", t->GetTemplateName()));
        }
        w.Open("DatablockTemplate");
        w.Text("Name", t->GetTemplateName());
        w.UInt("Stride", d.cbElementStride);
        w.Int("ParamType", (int)t->GetParameterBaseType());
        w.UInt("ElementsX", d.uiXElements);
        w.UInt("ElementsY", d.uiYElements);
        w.UInt("ElementsZ", d.uiZElements);
        w.Bool("IsRecordStream", t->DescribesRecordStream() == TRUE);
        w.Bool("IsByteAddressable", t->IsByteAddressable() == TRUE);
        if (d.cbPitch != d.cbElementStride * d.uiXElements) w.UInt("Pitch", d.cbPitch);
        for (UINT c = 1; c < NUM_DATABLOCK_CHANNELS; ++c) {
            BUFFERDIMENSIONS e = t->GetBufferDimensions(c);
            if (!e.uiXElements && !e.uiYElements && !e.uiZElements && !e.cbElementStride && !e.cbPitch) continue;
            w.Open("BufferDimensions");
            w.UInt("Channel", c);
            w.UInt("Stride", e.cbElementStride);
            w.UInt("ElementsX", e.uiXElements);
            w.UInt("ElementsY", e.uiYElements);
            w.UInt("ElementsZ", e.uiZElements);
            w.UInt("Pitch", e.cbPitch);
            w.Close();
        }
        if (t->HasInitialValue()) {
            w.Open("InitialValue");
            w.UInt("Elements", t->GetInitialValueElements());
            w.Bool("ExplicitlyEmpty", t->IsInitialValueEmpty() == TRUE);
            w.Text("Bytes", Hex(t->GetInitialValue(), t->GetInitialValueSizeBytes()));
            w.Close();
        }
        w.Close();
    }
    w.Close();

    w.Comment("  CompiledKernels (KernelID associates each Task with its CompiledKernel)  ");
    w.Open("CompiledKernels");
    w.UInt("Count", kernels.size());
    for (size_t k = 0; k < kernels.size(); ++k) {
        if (!strcasecmp(kernels[k]->GetSourceFile(), "HostFunction"))
            throw XmlError(Fmt("kernel %s was made from a function pointer and has no file to name", kernels[k]->GetOperation()));
        w.Open("CompiledKernel");
        w.Int("KernelID", (long long)k);
        w.Text("SourceFile", kernels[k]->GetSourceFile());
        w.Text("Operation", kernels[k]->GetOperation());
        w.Close();
    }
    w.Close();

    auto writePort = [&](Port *p, UINT index) {
        w.Open("Port");
        w.UInt("UniqueID", p->GetUID());
        w.UInt("OriginalIndex", index);
        w.Int("Type", (int)p->GetPortType());
        w.Text("DatablockTemplate", p->GetTemplate() ? p->GetTemplate()->GetTemplateName() : "");
        w.Text("VariableBinding", p->GetVariableBinding());
        w.UInt("ParameterIndex", p->GetFormalParameterIndex());
        w.UInt("InOutIndex", p->GetInOutRoutingIndex());
        if (p->GetMetaFunction() != MF_NONE) w.Int("MetaFunction", (int)p->GetMetaFunction());
        if (p->IsPermanentBlock()) w.Bool("PermanentBlock", true);
        if (!p->IsMarshallable()) w.Bool("Marshallable", false);
        w.Close();
    };
    w.Comment("  Tasks  ");
    w.Open("Tasks");
    w.UInt("Count", m_vTasks.size());
    for (Task *t : m_vTasks) {
        if (t->GetGeometryEstimator())
            throw XmlError(Fmt("task %s has a geometry estimator function, which has no text form",
                                        t->GetTaskName()));
        w.Open("Task");
        w.Text("Name", t->GetTaskName());
        w.Int("KernelID", kernelId[t->GetCompiledKernel()]);
        w.Int("AcceleratorClass", (int)t->GetAcceleratorClass());
        w.Open("InputPorts");
        w.UInt("Count", t->GetInputPortCount());
        for (UINT i = 0; i < t->GetInputPortCount(); ++i) writePort(t->GetInputPort(i), i);
        w.Close();
        w.Open("OutputPorts");
        w.UInt("Count", t->GetOutputPortCount());
        for (UINT i = 0; i < t->GetOutputPortCount(); ++i) writePort(t->GetOutputPort(i), i);
        w.Close();
        if (t->GetPriority()) w.Int("Priority", t->GetPriority());
        if (t->GetUserCodeAllocatesMemory()) w.Bool("UserCodeAllocatesMemory", true);
        w.Close();
    }
    w.Close();

    auto portUid = [&](Port *p) -> long long { return p ? (long long)p->GetUID() : -1; };
    auto isSwitch = [](Channel *c) {
        Port *d = c->GetBoundPort(CE_DST);
        if (!d) return false;
        for (UINT i = 0; i < d->GetControlChannelCount(); ++i)
            if (d->GetControlChannel(i) == c) return true;
        return false;
    };
    auto channelProps = [&](Channel *c) {
        w.UInt("Capacity", c->GetCapacity());
        if (c->GetViewMaterializationPolicy() != VIEWMATERIALIZATIONPOLICY_ON_DEMAND)
            w.Int("ViewMaterialization", (int)c->GetViewMaterializationPolicy());
        if (!c->ShouldDraw()) w.Bool("Draw", false);
    };
    w.Comment("  Channels (ports are referred to by their UniqueID)  ");
    w.Open("Channels");
    w.UInt("Count", m_vChannels.size() - members.size());
    for (Channel *c : m_vChannels) {
        if (members.count(c)) continue;  // written inside their multi-channel
        Port *src = c->GetBoundPort(CE_SRC), *dst = c->GetBoundPort(CE_DST);
        if ((src && !known(src)) || (dst && !known(dst)))
            throw XmlError(Fmt("channel %s is bound to a port of no task in this graph", c->GetName()));
        w.Open("Channel");
        w.Text("Name", c->GetName());
        w.Int("Type", (int)c->GetType());
        w.Int("SourcePort", portUid(src));
        w.Int("DestinationPort", portUid(dst));
        if (isSwitch(c)) w.Bool("Switch", true);
        channelProps(c);
        if (c->GetType() == CT_INITIALIZER) {
            CONTROLSIGNAL s = static_cast<InitializerChannel *>(c)->GetInitialPropagatedControlSignal();
            if (s != DBCTLC_NONE) w.UInt("InitialControlSignal", s);
        }
        if (c->GetType() == CT_MULTI) {
            std::vector<Channel *> &mem = static_cast<MultiChannel *>(c)->GetMembers();
            bool sw = !mem.empty() && isSwitch(mem[0]);
            for (Channel *m : mem) {
                if (isSwitch(m) != sw)
                    throw XmlError(Fmt("multi-channel %s mixes switch and normal members", c->GetName()));
                if (!known(m->GetBoundPort(CE_DST)))
                    throw XmlError(Fmt("multi-channel %s feeds a port of no task in this graph", c->GetName()));
            }
            if (sw) w.Bool("Switch", true);
            w.Open("Members");
            w.UInt("Count", mem.size());
            for (Channel *m : mem) {
                w.Open("Member");
                w.Text("Name", m->GetName());
                w.Int("DestinationPort", portUid(m->GetBoundPort(CE_DST)));
                channelProps(m);
                w.Close();
            }
            w.Close();
        }
        w.Close();
    }
    w.Close();

    // Actions: the bindings, read off the ports in task order, then the geometry of every task
    // and the predicates of every channel endpoint (reference: Recorder.cpp Write methods).
    XmlOut a;  // collected first: <Count> precedes them
    size_t nActions = 0;
    auto action = [&](const char *name) {
        a.Open("Action");
        a.Text("ActionName", name);
        ++nActions;
    };
    for (Port *p : ports) {
        for (Port *g : p->GetGatedPorts()) {
            if (!known(g)) throw XmlError(Fmt("port %s gates a port of no task", p->GetVariableBinding()));
            action("BindControlPort");
            a.UInt("ControllerPort", p->GetUID());
            a.UInt("GatedPort", g->GetUID());
            a.Bool("InitiallyOpen", static_cast<OutputPort *>(g)->IsInitiallyOpen() == TRUE);
            a.Close();
        }
        for (Port *o : p->GetControlPropagationPorts()) {
            if (!known(o)) throw XmlError(Fmt("port %s propagates to a port of no task", p->GetVariableBinding()));
            action("BindControlPropagationPort");
            a.UInt("InputPort", p->GetUID());
            a.UInt("OutputPort", o->GetUID());
            a.Close();
        }
        for (Channel *c : p->GetControlPropagationChannels()) {
            if (!c || !GetChannel(c->GetName()))
                throw XmlError(Fmt("port %s propagates to a channel of another graph", p->GetVariableBinding()));
            action("BindControlPropagationChannel");
            a.UInt("InputPort", p->GetUID());
            a.Text("ControlledChannel", c->GetName());
            a.Close();
        }
        for (auto &d : p->GetDescriptorPorts()) {
            if (!known(d.first)) throw XmlError(Fmt("port %s is described by a port of no task", p->GetVariableBinding()));
            action("BindDescriptorPort");
            a.UInt("DescribedPort", p->GetUID());
            a.UInt("DescriberPort", d.first->GetUID());
            a.Int("DescriptorFunction", (int)d.second);
            a.Close();
        }
        for (Port *s : p->GetIterationTargets()) {
            if (!known(s)) throw XmlError(Fmt("port %s scopes a port of no task", p->GetVariableBinding()));
            action("BindIterationScope");
            a.UInt("MetaPort", p->GetUID());
            a.UInt("ScopedPort", s->GetUID());
            a.Close();
        }
    }
    for (Task *t : m_vTasks) {
        int x, y, z;
        t->GetComputeGeometry(&x, &y, &z);
        action("SetComputeGeometry");
        a.Text("TaskName", t->GetTaskName());
        a.Int("PreferredXDim", x);
        a.Int("PreferredYDim", y);
        a.Int("PreferredZDim", z);
        a.Close();
        PTASKDIM3 grid, block;
        if (t->GetBlockAndGridSize(&grid, &block)) {
            action("SetBlockAndGridSize");
            a.Text("TaskName", t->GetTaskName());
            a.Int("BlockXDim", block.x);
            a.Int("BlockYDim", block.y);
            a.Int("BlockZDim", block.z);
            a.Int("GridXDim", grid.x);
            a.Int("GridYDim", grid.y);
            a.Int("GridZDim", grid.z);
            a.Close();
        }
    }
    for (Channel *c : m_vChannels) {
        for (CHANNELENDPOINTTYPE e : {CE_SRC, CE_DST}) {
            CHANNELPREDICATE pr = c->GetPredicationType(e);
            if (pr == CGATEFN_NONE) continue;
            if (pr == CGATEFN_USER_DEFINED)
                throw XmlError(Fmt("channel %s has a user-defined predicate function, which has no text form", c->GetName()));
            action("SetPredicationType");
            a.Text("ChannelName", c->GetName());
            a.Int("EndpointType", (int)e);
            a.Int("CanonicalPredicate", (int)pr);
            a.Int("FailureAction", (int)c->GetPredicateFailureAction(e));
            a.Close();
        }
    }
    w.Open("Actions");
    w.UInt("Count", nActions);
    std::string doc = w.Str();
    // splice the actions in at the current depth (two levels: Graph, Actions)
    const std::string &acts = a.Str();
    size_t body = acts.find('\n') + 1;  // past the declaration the second writer emitted
    for (size_t i = body; i < acts.size();) {
        size_t e = acts.find('\n', i);
        doc += "    " + acts.substr(i, e - i + 1);
        i = e + 1;
    }
    doc += "  </Actions>\n</Graph>\n";
    return doc;
}

// reference: Graph::Serialize graph.cpp:6452-6468 and XMLWriter::WriteGraph XMLWriter.cpp:124-245
BOOL Graph::Serialize(const char *lpszFileName) {
    const std::string where = std::string("Serialize(") + (lpszFileName ? lpszFileName : "") + "): ";
    if (!lpszFileName || !*lpszFileName) return Runtime::HandleError("Serialize: no file name");
    std::string doc;
    try {
        // the graph lock is dropped before HandleError runs
        std::lock_guard<std::recursive_mutex> lk(m_lock);
        doc = SerializeLocked();
    } catch (const XmlError &e) {
        return Runtime::HandleError("%s%s", where.c_str(), e.what());
    }

    std::string tmp = std::string(lpszFileName) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return Runtime::HandleError("%scannot open %s for writing: %s", where.c_str(), tmp.c_str(), strerror(errno));
    bool ok = fwrite(doc.data(), 1, doc.size(), f) == doc.size();
    ok = fclose(f) == 0 && ok;
    if (!ok || rename(tmp.c_str(), lpszFileName) != 0) {
        remove(tmp.c_str());
        return Runtime::HandleError("%swriting the file failed: %s", where.c_str(), strerror(errno));
    }
    return TRUE;
}

// =============================================================================================
// Graph::Deserialize
// =============================================================================================

// reference: Graph::Deserialize graph.cpp:6477-6490 and XMLReader::ReadGraph XMLReader.cpp:383-404
// (templates, kernels, tasks, channels, then the recorded actions, in that order)
BOOL Graph::Deserialize(const char *lpszFileName) {
    const std::string where = std::string("Deserialize(") + (lpszFileName ? lpszFileName : "") + "): ";
    if (!m_vTasks.empty() || !m_vChannels.empty())
        return Runtime::HandleError("%sthe graph already has tasks or channels", where.c_str());
    std::string text;
    {
        FILE *f = lpszFileName ? fopen(lpszFileName, "rb") : nullptr;
        if (!f) return Runtime::HandleError("%scannot open the file: %s", where.c_str(), strerror(errno));
        char buf[65536];
        size_t n;
        while ((n = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
        fclose(f);
    }
    try {
        XmlNode root;
        XmlParser(text).Document(&root);
        if (root.tag != "Graph") throw XmlError("the root element is <" + root.tag + ">, not <Graph>");
        if (root.Find("Name") && m_bUnnamed) m_strName = root.Str("Name");

        // ---- templates: registered with the runtime; one of the same name must match
        std::map<std::string, DatablockTemplate *> templates;
        for (const XmlNode *n : root.Need("DatablockTemplates").List("DatablockTemplate")) {
            std::string name = n->Str("Name");
            UINT stride = n->Uint32("Stride");
            PTASK_PARM_TYPE type = (PTASK_PARM_TYPE)n->Int("ParamType");
            UINT x = n->Uint32("ElementsX"), y = n->Uint32("ElementsY"), z = n->Uint32("ElementsZ");
            bool rec = n->Bool("IsRecordStream"), raw = n->Bool("IsByteAddressable");
            UINT pitch = (UINT)n->UIntOr("Pitch", (unsigned long long)stride * x);
            DatablockTemplate *t = nullptr;
            if (IsScalarType(type)) {
                t = new DatablockTemplate(name.c_str(), stride, type);
            } else {
                BUFFERDIMENSIONS dims[NUM_DATABLOCK_CHANNELS];
                dims[DBDATA_IDX].Initialize(x, y, z, stride, pitch);
                for (UINT c = 1; c < NUM_DATABLOCK_CHANNELS; ++c) dims[c].Initialize(0, 0, 0, 0, 0);
                for (const XmlNode *b : n->All("BufferDimensions")) {
                    UINT c = b->Uint32("Channel");
                    if (c == DBDATA_IDX || c >= NUM_DATABLOCK_CHANNELS) throw XmlError("template " + name + ": bad BufferDimensions channel");
                    dims[c].Initialize(b->Uint32("ElementsX"), b->Uint32("ElementsY"), b->Uint32("ElementsZ"), b->Uint32("Stride"),
                                       b->Uint32("Pitch"));
                }
                t = new DatablockTemplate(name.c_str(), dims, NUM_DATABLOCK_CHANNELS, rec, raw);
            }
            if (const XmlNode *iv = n->Find("InitialValue")) {
                std::vector<unsigned char> bytes = Unhex(iv->Str("Bytes"));
                t->SetInitialValue(bytes.empty() ? nullptr : bytes.data(), (UINT)bytes.size(), iv->Uint32("Elements"),
                                   iv->BoolOr("ExplicitlyEmpty", false) ? TRUE : FALSE);
            }
            DatablockTemplate *have = Runtime::GetDatablockTemplate((char *)name.c_str());
            if (have) {
                bool same = TemplateSignature(have) == TemplateSignature(t);
                delete t;
                if (!same) throw XmlError("template " + name + " differs from the runtime's template of that name");
                t = have;
            } else {
                // registered through the runtime's table like every other template
                DatablockTemplate *reg = Runtime::RegisterTemplate(t);
                if (reg != t) throw XmlError("template " + name + " could not be registered");
            }
            templates[name] = t;
        }

        // ---- kernels
        std::map<long long, CompiledKernel *> kernels;
        for (const XmlNode *n : root.Need("CompiledKernels").List("CompiledKernel")) {
            std::string file = n->Str("SourceFile"), op = n->Str("Operation");
            ACCELERATOR_CLASS cls = ptaskutils::SelectAcceleratorClass(file.c_str());
            if (cls == ACCELERATOR_CLASS_CUDA) {
                // a graph written for CUDA: its kernels are looked for as the same-stem code object
                size_t dot = file.rfind('.');
                file = file.substr(0, dot) + ".hsaco";
            }
            char log[1024] = {0};
            CompiledKernel *k = Runtime::GetCompiledKernel((char *)file.c_str(), (char *)op.c_str(), log, sizeof(log));
            if (!k) throw XmlError("kernel " + op + " of " + file + " could not be loaded" + (log[0] ? std::string(": ") + log : ""));
            long long id = n->Int("KernelID");
            if (kernels.count(id)) throw XmlError("KernelID " + std::to_string(id) + " appears twice");
            kernels[id] = k;
        }

        // ---- tasks and their ports
        std::map<UINT, Port *> ports;
        auto readPorts = [&](const XmlNode &list, std::vector<Port *> *out) {
            std::vector<const XmlNode *> items = list.List("Port");
            out->assign(items.size(), nullptr);
            for (const XmlNode *n : items) {
                UINT idx = n->Uint32("OriginalIndex");
                if (idx >= items.size() || (*out)[idx]) throw XmlError("port OriginalIndex " + std::to_string(idx) + " is out of place");
                std::string tn = n->Str("DatablockTemplate");
                auto it = templates.find(tn);
                if (it == templates.end()) throw XmlError("port names unknown template '" + tn + "'");
                UINT uid = n->Uint32("UniqueID");
                if (ports.count(uid)) throw XmlError("port UniqueID " + std::to_string(uid) + " appears twice");
                std::string binding = n->Str("VariableBinding");
                Port *p = Runtime::CreatePort((PORTTYPE)n->Int("Type"), it->second, uid, (char *)binding.c_str(),
                                              n->Uint32("ParameterIndex"), n->Uint32("InOutIndex"));
                if (!p) throw XmlError("port " + binding + " could not be created");
                p->SetMetaFunction((METAFUNCTION)n->IntOr("MetaFunction", MF_NONE));
                if (n->BoolOr("PermanentBlock", false)) p->SetPermanentBlock(TRUE);
                if (!n->BoolOr("Marshallable", true)) p->SetMarshallable(FALSE);
                (*out)[idx] = p;
                ports[uid] = p;
            }
        };
        for (const XmlNode *n : root.Need("Tasks").List("Task")) {
            std::string name = n->Str("Name");
            auto k = kernels.find(n->Int("KernelID"));
            if (k == kernels.end()) throw XmlError("task " + name + " names an unknown KernelID");
            std::vector<Port *> ins, outs;
            readPorts(n->Need("InputPorts"), &ins);
            readPorts(n->Need("OutputPorts"), &outs);
            ACCELERATOR_CLASS recorded = (ACCELERATOR_CLASS)n->Int("AcceleratorClass");
            ACCELERATOR_CLASS cls = k->second->GetAcceleratorClass();
            if (recorded != cls && !(recorded == ACCELERATOR_CLASS_CUDA && cls == ACCELERATOR_CLASS_HIP))
                throw XmlError("task " + name + " was recorded for accelerator class " + std::to_string((int)recorded) +
                                " but its kernel loads for class " + std::to_string((int)cls));
            Task *t = AddTask(k->second, (UINT)ins.size(), ins.data(), (UINT)outs.size(), outs.data(), name.c_str());
            if (!t) throw XmlError("task " + name + " could not be added");
            if (n->Find("Priority")) t->SetPriority((int)n->Int("Priority"));
            if (n->BoolOr("UserCodeAllocatesMemory", false)) t->SetUserCodeAllocatesMemory(TRUE);
        }
        auto port = [&](const XmlNode &n, const char *tag) -> Port * {
            long long uid = n.Int(tag);
            auto it = ports.find((UINT)uid);
            if (uid < 0 || it == ports.end()) throw XmlError("<" + std::string(tag) + "> names unknown port " + std::to_string(uid));
            return it->second;
        };

        // ---- channels
        auto props = [](const XmlNode &n, Channel *c) {
            if (n.Find("Capacity")) c->SetCapacity(n.Uint32("Capacity"));
            if (n.Find("ViewMaterialization")) c->SetViewMaterializationPolicy((VIEWMATERIALIZATIONPOLICY)n.Int("ViewMaterialization"));
            if (!n.BoolOr("Draw", true)) c->SetNoDraw();
        };
        for (const XmlNode *n : root.Need("Channels").List("Channel")) {
            std::string name = n->Str("Name");
            if (GetChannel(name.c_str())) throw XmlError("channel " + name + " appears twice");
            BOOL sw = n->BoolOr("Switch", false) ? TRUE : FALSE;
            Channel *c = nullptr;
            switch ((CHANNELTYPE)n->Int("Type")) {
                case CT_GRAPH_INPUT: c = AddInputChannel(port(*n, "DestinationPort"), name.c_str(), sw); break;
                case CT_INITIALIZER: {
                    InitializerChannel *ic = AddInitializerChannel(port(*n, "DestinationPort"), name.c_str(), sw);
                    if (n->Find("InitialControlSignal")) ic->SetInitialPropagatedControlSignal(n->UInt("InitialControlSignal"));
                    c = ic;
                    break;
                }
                case CT_GRAPH_OUTPUT: c = AddOutputChannel(port(*n, "SourcePort"), name.c_str()); break;
                case CT_INTERNAL:
                    c = AddInternalChannel(port(*n, "SourcePort"), port(*n, "DestinationPort"), name.c_str(), sw);
                    break;
                case CT_MULTI: {
                    std::vector<const XmlNode *> mem = n->Need("Members").List("Member");
                    std::vector<Port *> dst;
                    for (const XmlNode *m : mem) dst.push_back(port(*m, "DestinationPort"));
                    MultiChannel *mc = AddMultiChannel(dst.data(), (int)dst.size(), name.c_str(), sw);
                    for (size_t i = 0; i < mem.size(); ++i) {
                        Channel *m = mc->GetMembers()[i];
                        if (mem[i]->Str("Name") != m->GetName())
                            throw XmlError("multi-channel " + name + " member " + std::to_string(i) + " is named '" + m->GetName() +
                                            "' here, not '" + mem[i]->Str("Name") + "'");
                        props(*mem[i], m);
                    }
                    c = mc;
                    break;
                }
                default: throw XmlError("channel " + name + " has unknown Type " + n->Str("Type"));
            }
            if (!c) throw XmlError("channel " + name + " could not be added");
            props(*n, c);
        }

        // ---- actions, in document order
        auto channel = [&](const XmlNode &n, const char *tag) -> Channel * {
            Channel *c = GetChannel(n.Str(tag).c_str());
            if (!c) throw XmlError("<" + std::string(tag) + "> names unknown channel '" + n.Str(tag) + "'");
            return c;
        };
        auto task = [&](const XmlNode &n) -> Task * {
            Task *t = GetTask(n.Str("TaskName").c_str());
            if (!t) throw XmlError("unknown task '" + n.Str("TaskName") + "'");
            return t;
        };
        if (const XmlNode *acts = root.Find("Actions")) {
            for (const XmlNode *n : acts->List("Action")) {
                std::string an = n->Str("ActionName");
                if (an == "BindControlPort") {
                    Port *g = port(*n, "GatedPort");
                    if (g->GetPortType() != OUTPUT_PORT) throw XmlError("BindControlPort gates a port that is not an output");
                    BindControlPort(port(*n, "ControllerPort"), g, n->Bool("InitiallyOpen") ? TRUE : FALSE);
                } else if (an == "BindControlPropagationPort") {
                    BindControlPropagationPort(port(*n, "InputPort"), port(*n, "OutputPort"));
                } else if (an == "BindControlPropagationChannel") {
                    BindControlPropagationChannel(port(*n, "InputPort"), channel(*n, "ControlledChannel"));
                } else if (an == "BindDescriptorPort") {
                    BindDescriptorPort(port(*n, "DescribedPort"), port(*n, "DescriberPort"), (DESCRIPTORFUNC)n->Int("DescriptorFunction"));
                } else if (an == "BindIterationScope") {
                    BindIterationScope(port(*n, "MetaPort"), port(*n, "ScopedPort"));
                } else if (an == "SetComputeGeometry") {
                    task(*n)->SetComputeGeometry((int)n->Int("PreferredXDim"), (int)n->Int("PreferredYDim"), (int)n->Int("PreferredZDim"));
                } else if (an == "SetBlockAndGridSize") {
                    task(*n)->SetBlockAndGridSize(
                        PTASKDIM3((int)n->Int("GridXDim"), (int)n->Int("GridYDim"), (int)n->Int("GridZDim")),
                        PTASKDIM3((int)n->Int("BlockXDim"), (int)n->Int("BlockYDim"), (int)n->Int("BlockZDim")));
                } else if (an == "SetPredicationType") {
                    long long e = n->Int("EndpointType");
                    if (e != CE_SRC && e != CE_DST) throw XmlError("SetPredicationType: bad EndpointType");
                    long long pr = n->Int("CanonicalPredicate");
                    if (pr < CGATEFN_NONE || pr >= CGATEFN_USER_DEFINED) throw XmlError("SetPredicationType: bad CanonicalPredicate");
                    channel(*n, "ChannelName")
                        ->SetPredicationType((CHANNELENDPOINTTYPE)e, (CHANNELPREDICATE)pr,
                                             (PREDICATE_FAILURE_ACTION)n->IntOr("FailureAction", PFA_RELEASE_BLOCK));
                } else {
                    throw XmlError("unknown action '" + an + "'");
                }
            }
        }
    } catch (const XmlError &e) {
        return Runtime::HandleError("%s%s", where.c_str(), e.what());
    }
    return TRUE;
}

}  // namespace PTask
