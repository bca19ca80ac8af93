// Marching-cubes lookup tables shared by the HIP kernels, the C++ host
// mirror and the CPU oracle.
//
// The triangle table is the standard public Lorensen/Bourke marching-cubes
// triangulation ("Polygonising a scalar field", P. Bourke, 1994), which is
// also what the reference crate ships (src/marching_cubes_tables.rs:74-331,
// TRIANGLE_CONNECTION).  Topology parity with the reference requires the
// identical triangulation, so the data is the same; the encoding is ours:
// case c lists its triangles as a string of hex edge ids, three per
// triangle, in table order.  The test suite pins the decoded int8[256][16]
// table (row-major, -1 padded) to SHA-256
//   19bf7699e214903d72c94c296546f2e31337d637a1e4b118c3108a0f428e809b
// and the derived edge mask table (u16le[256]) to
//   2572f4b459bce46d930e57b2c261ef8143b6d43157272c698c64d70dd295b596
// (both computed from the reference's own table text).
//
// Everything else here is derived at compile time from that one table:
// the crossed-edge mask per case (reference EDGE_CROSSING_MASK,
// marching_cubes_tables.rs:49-70, equals the set of edges the case's
// triangles use), the triangle count per case, and for each case and edge
// the set of edges whose first appearance in the triangle list precedes it
// (used to reproduce the reference's first-encounter vertex numbering,
// mesh.rs:240-251).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define ISO_HD __host__ __device__
#else
#define ISO_HD
#endif

namespace iso {
namespace tables {

// Corner offsets of a cell, reference CORNERS (marching_cubes_tables.rs:16-25).
inline constexpr int kCorner[8][3] = {
    {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// Corner pair of each cell edge, reference EDGE_CONNECTION (:32-45).
inline constexpr int kEdgeCorners[12][2] = {
    {0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
    {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// Morton child order -> cell corner order, reference REMAP_CUBE (:29).
inline constexpr int kRemapCube[8] = {2, 3, 1, 0, 6, 7, 5, 4};

inline constexpr const char* kTriEncoded[256] = {
    "", "083", "019", "183981", "12a", "08312a", "92a029", "2832a8a98",
    "3b2", "0b28b0", "19023b", "1b219b98b", "3a1ba3", "0a108a8ba", "3903b9ba9", "98aa8b",
    "478", "430734", "019847", "419471731", "12a847", "34730412a", "92a902847", "2a9297273794",
    "8473b2", "b47b24204", "90184723b", "47b94b9b2921", "3a13ba784", "1ba14b1047b4", "47890b9bab03", "47b4b99ba",
    "954", "954083", "054150", "854835315", "12a954", "30812a495", "52a542402", "2a5325354348",
    "95423b", "0b208b495", "05401523b", "21525828b485", "a3ba13954", "4950818a18ba", "54050b5bab03", "54858aa8b",
    "978579", "930953573", "078017157", "153357", "978957a12", "a12950530573", "802825857a52", "2a5253357",
    "7957893b2", "95797292027b", "23b018178157", "b21b17715", "958857a13a3b", "5705097b010aba0", "ba0b03a50807570", "ba57b5",
    "a65", "0835a6", "9015a6", "1831985a6", "165261", "165126308", "965906026", "598582526328",
    "23ba65", "b08b20a65", "01923b5a6", "5a61929b298b", "63b653513", "08b0b50515b6", "3b6036065059", "65969bb98",
    "5a6478", "43047365a", "1905a6847", "a65197173794", "612651478", "125526304347", "847905065026", "739794329596269",
    "3b2784a65", "5a647242027b", "01947823b5a6", "9219b294b7b45a6", "8473b53515b6", "51b5b610b7b404b", "059065036b63847", "65969b4797b9",
    "a4964a", "4a649a083", "a01a60640", "83181686461a", "149124264", "308129249264", "024426", "832824426",
    "a49a64b23", "08228b49a4a6", "3b201606461a", "64161a48121b8b1", "964936913b63", "8b1810b61914641", "3b6360064", "648b68",
    "7a678a89a", "0730a709a67a", "a671a7178180", "a67a71173", "126168189867", "269291679093739", "780706602", "732672",
    "23ba68a89867", "20727b09767a9a7", "1801781a767a23b", "b21b17a61671", "896867916b63136", "091b67", "7807063b0b60", "7b6",
    "76b", "308b76", "019b76", "819831b76", "a126b7", "12a3086b7", "2902a96b7", "6b72a3a83a98",
    "723627", "708760620", "276237019", "162186198876", "a76a17137", "a7617a187108", "03707a0a96a7", "76a7a88a9",
    "684b86", "36b306046", "86b846901", "946963931b36", "6846b82a1", "12a30b06b046", "4b846b0292a9", "a93a32943b36463",
    "823842462", "042462", "190234246438", "194142246", "8138618466a1", "a10a06604", "4634386a3039a93", "a946a4",
    "49576b", "083495b76", "50154076b", "b76834354315", "954a1276b", "6b712a083495", "76b54a42a402", "348354325a52b76",
    "723762549", "954086062687", "362376150540", "628687218485158", "954a16176137", "16a176107870954", "40a4a503a6a737a", "76a7a854a48a",
    "6956b9b89", "36b063056095", "0b805b01556b", "6b3635531", "12a95b9b8b56", "0b306b09656912a", "b85b56805a52025", "6b36352a3a53",
    "589528562382", "956960062", "158180568382628", "156216", "13616a386569896", "a10a06950560", "03856a", "a56",
    "b5a75b", "b5ab75830", "5b75ab190", "a75ab7981831", "b12b71751", "08312717572b", "9759279022b7", "75272b592328982",
    "25a235375", "820852875a25", "9015a35373a2", "982921872a25752", "135375", "087071175", "903935537", "987597",
    "5845a8ab8", "5045b05abb30", "01984a8aba45", "ab4a45b34941314", "2512852b8458", "04b0b345b2b151b", "0250592b5458b85", "9452b3",
    "25a352345384", "5a2524420", "3a235a385458019", "5a2524192942", "845853351", "045105", "845853905035", "945",
    "4b749b9ab", "0834979b79ab", "1ab1b414074b", "3143481a474bab4", "4b79b492b912", "9749b791b2b1083", "b74b42240", "b74b42834324",
    "29a279237749", "9a7974a27870207", "37a3a274a1a040a", "1a2874", "491417713", "491417081871", "403743", "487",
    "9a8ab8", "30939bb9a", "01a0a88ab", "31ab3a", "12b1b99b8", "30939b1292b9", "02b80b", "32b",
    "23828aa89", "9a2092", "23828a0181a8", "1a2", "138918", "091", "038", "",
};

struct McTables {
  int8_t tri[256][16];        // triangle edge ids, -1 padded
  uint16_t edge_mask[256];    // edges crossed by the case
  uint8_t tri_count[256];     // triangles emitted by the case
  uint16_t before[256][12];   // edges first used before edge e (by case)
};

constexpr int hex_digit(char c) {
  return (c >= '0' && c <= '9') ? c - '0' : c - 'a' + 10;
}

constexpr McTables make_tables() {
  McTables t{};
  for (int c = 0; c < 256; ++c) {
    for (int k = 0; k < 16; ++k) t.tri[c][k] = -1;
    const char* s = kTriEncoded[c];
    int n = 0;
    uint16_t seen = 0;
    uint16_t order[12] = {};
    int n_order = 0;
    while (s[n] != '\0') {
      int e = hex_digit(s[n]);
      t.tri[c][n] = static_cast<int8_t>(e);
      if (!(seen & (1u << e))) {
        seen = static_cast<uint16_t>(seen | (1u << e));
        order[n_order++] = static_cast<uint16_t>(e);
      }
      ++n;
    }
    t.edge_mask[c] = seen;
    t.tri_count[c] = static_cast<uint8_t>(n / 3);
    for (int e = 0; e < 12; ++e) t.before[c][e] = 0;
    uint16_t acc = 0;
    for (int k = 0; k < n_order; ++k) {
      t.before[c][order[k]] = acc;
      acc = static_cast<uint16_t>(acc | (1u << order[k]));
    }
  }
  return t;
}

inline constexpr McTables kTables = make_tables();

// The polygon loops of each case that ExtendedMarchingCubes fans around a
// feature vertex (reference EDGE_LOOPS, marching_cubes_tables.rs:337-594,
// read by extended_marching_cubes.rs:135-155 as: loop count, the loops'
// sizes, then their edges).  Same data, our encoding: per case the loop
// count, then per loop its size and its edges, as hex digits.  Three rows of
// the reference's text carry entries after the last loop that its code never
// reads (cases 167, 181, 229); they are not kept.  27 cases list an edge in
// two loops (a hexagon split into two quads along a diagonal): the vertex is
// shared and the diagonal is fanned twice.  The test suite pins the decoded
// table against SHA-256 of the reference's text.
inline constexpr const char* kLoopsEncoded[256] = {
    "0", "13083", "13019", "141983", "1312a", "23083312a",
    "142a90", "15a9832", "133b2", "1408b2", "23190323b", "1598b21",
    "143ba1", "158ba10", "15ba903", "148ba9", "13478", "144730",
    "230193847", "1573194", "2312a3847", "243047312a", "242a903847", "16732a94",
    "2384733b2", "152047b", "339013847323b", "24219b4b947", "243ba13784", "161047ba",
    "23478503ba9", "1547ba9", "13954", "239543083", "140154", "1531548",
    "2312a3954", "33308312a3495", "15402a5", "242a5344835", "23954323b", "2408b23495",
    "240154323b", "1621548b", "243ba13954", "234955108ba", "165403ba", "15548ba",
    "147895", "1557309", "1515780", "143157", "2478953a12", "23a12509573",
    "162a5780", "152a573", "24789533b2", "1620957b", "2323b580157", "15b2157",
    "243ba145789", "1757ba109", "17ba57803", "1457ba", "13a65", "2308335a6",
    "2390135a6", "24198335a6", "141265", "2412653308", "1502659", "16265983",
    "2323b3a65", "2408b23a65", "33019323b35a6", "235a652198b", "15513b6", "165108b6",
    "243b6045906", "156598b", "235a63478", "244730365a", "3319035a63847", "23a65594731",
    "2412653478", "24265143047", "23847559026", "177326594", "333b237843a65", "235a657b204",
    "430193478323b35a6", "34219b447b935a6", "238475b6513", "1751047b6", "34065943b603847", "24947b4659b",
    "1449a6", "2449a63083", "156401a", "161a6483", "1526491", "23308591264",
    "142640", "1583264", "24a6493b23", "242b804a649", "233b251a640", "17648b21a",
    "163b6491", "178b64910", "153b640", "148b64", "1589a67", "1609a673",
    "2480174a671", "15a6731", "16126789", "172673091", "1578026", "147326",
    "2323b56789a", "17209a67b", "3480174a6713b23", "24b21741a67", "178913b67", "230913b67",
    "2403b647806", "137b6", "1376b", "233083b76", "230193b76", "2419833b76",
    "23a1236b7", "3312a330836b7", "242a9036b7", "236b7532a98", "142376", "1562087",
    "2423763019", "16621987", "151376a", "24a17648710", "16a90376", "1576a98",
    "146b84", "15046b3", "246b843901", "166b3194", "246b8432a1", "2312a5b3046",
    "2446b842a90", "17a946b32", "1546238", "144620", "23190538462", "1519462",
    "1613846a", "15a1046", "1746a9038", "1446a9", "23495376b", "3308334953b76",
    "240154376b", "23b76548315", "339543a12376b", "436b7312a30833495", "2376b5a5402", "34532a4483536b7",
    "2423763549", "23954587620", "24376240154", "176215487", "2395456a137", "34087146a173954",
    "1740376a5", "2448a5476a8", "15b8956", "24095646b30", "160156b8", "156b315",
    "2312a556b89", "34b3064956032a1", "17b802a56", "246b354a532", "16238956", "1595620",
    "171562380", "141562", "17138956a", "2456094a106", "23038356a", "13a56",
    "145ab7", "245ab73830", "245ab73190", "24ab7541983", "157512b", "2308352b751",
    "162b7590", "17759832b", "15375a2", "165a2087", "239015a2375", "179875a21",
    "143751", "1508751", "1590375", "147598", "15ab845", "16045ab3",
    "23019545ab8", "17ab31945", "162b8451", "1704512b3", "1702b8459", "2394532b3",
    "24235a44538", "155a204", "3435a2484533019", "16a21945", "1584513", "140451",
    "24035948453", "13945", "159ab74", "230835749ab", "161ab740", "1731ab748",
    "242b91449b7", "3412b94749b3830", "15b7402", "24b74243248", "1623749a", "179a20874",
    "1737401a2", "231a23874", "1549137", "24871044917", "143740", "13487",
    "1489ab", "15309ab", "1501ab8", "1431ab", "1512b89", "242b914309b",
    "1402b8", "1332b", "152389a", "14209a", "24238a41a80", "131a2",
    "141389", "13091", "13038", "0",
};

struct LoopTables {
  uint8_t count[256];        // loops of the case
  uint8_t size[256][4];      // vertices of each loop
  uint8_t total[256];        // sum of the sizes: the fan triangles of the case
  int8_t edge[256][12];      // loop edges, loop after loop, -1 padded
  uint16_t before[256][12];  // edges first listed before edge e (by case)
};

constexpr LoopTables make_loop_tables() {
  LoopTables t{};
  for (int c = 0; c < 256; ++c) {
    const char* s = kLoopsEncoded[c];
    for (int k = 0; k < 12; ++k) t.edge[c][k] = -1;
    for (int k = 0; k < 4; ++k) t.size[c][k] = 0;
    for (int e = 0; e < 12; ++e) t.before[c][e] = 0;
    int n = 0;
    const int loops = hex_digit(s[n++]);
    t.count[c] = static_cast<uint8_t>(loops);
    int m = 0;
    uint16_t seen = 0;
    for (int l = 0; l < loops; ++l) {
      const int size = hex_digit(s[n++]);
      t.size[c][l] = static_cast<uint8_t>(size);
      for (int k = 0; k < size; ++k) {
        const int e = hex_digit(s[n++]);
        t.edge[c][m++] = static_cast<int8_t>(e);
        if (!(seen & (1u << e))) {
          t.before[c][e] = seen;
          seen = static_cast<uint16_t>(seen | (1u << e));
        }
      }
    }
    t.total[c] = static_cast<uint8_t>(m);
  }
  return t;
}

inline constexpr LoopTables kLoops = make_loop_tables();

}  // namespace tables
}  // namespace iso
