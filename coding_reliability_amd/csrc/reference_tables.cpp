// Golden distance distributions the reference's result writers compare
// against (distance_distribution/src/solutions.h:18-114): Extended Hamming
// weight distributions for n = 8, 16, 24, 32 and exhaustive AN counts for
// A = 61, 641 at k = 8, 16, 24.  Data only; every value fits in 64 bits.
#include "reference_tables.h"

namespace cr {
namespace tables {

const unsigned long long sol_hamming_n8[14] = {
    0ull, 0ull, 0ull, 0ull, 55ull, 0ull, 96ull, 0ull, 87ull, 0ull, 16ull, 0ull, 1ull, 0ull,
};
const unsigned sol_hamming_n8_len = 14;

const unsigned long long sol_hamming_n16[23] = {
    0ull, 0ull, 0ull, 0ull, 263ull, 0ull, 2224ull, 0ull, 10202ull, 0ull, 19952ull, 0ull, 20414ull,
    0ull, 9872ull, 0ull, 2389ull, 0ull, 208ull, 0ull, 11ull, 0ull, 0ull,
};
const unsigned sol_hamming_n16_len = 23;

const unsigned long long sol_hamming_n24[31] = {
    0ull, 0ull, 0ull, 0ull, 945ull, 0ull, 18200ull, 0ull, 183885ull, 0ull, 936936ull, 0ull, 2705885ull,
    0ull, 4541040ull, 0ull, 4547475ull, 0ull, 2700880ull, 0ull, 939939ull, 0ull, 182520ull, 0ull,
    18655ull, 0ull, 840ull, 0ull, 15ull, 0ull, 0ull,
};
const unsigned sol_hamming_n24_len = 31;

const unsigned long long sol_hamming_n32[40] = {
    0ull, 0ull, 0ull, 0ull, 1583ull, 0ull, 51744ull, 0ull, 965668ull, 0ull, 9908896ull, 0ull,
    61116300ull, 0ull, 235727520ull, 0ull, 589304574ull, 0ull, 974127392ull, 0ull, 1077087634ull, 0ull,
    797231712ull, 0ull, 392806740ull, 0ull, 126854112ull, 0ull, 26224444ull, 0ull, 3311840ull, 0ull,
    238729ull, 0ull, 8288ull, 0ull, 119ull, 0ull, 0ull, 0ull,
};
const unsigned sol_hamming_n32_len = 40;

const unsigned long long solution_an8_A61[15] = {
    256ull, 0ull, 0ull, 1362ull, 4106ull, 8358ull, 12092ull, 13468ull, 11884ull, 8176ull, 4134ull,
    1262ull, 304ull, 134ull, 0ull,
};
const unsigned solution_an8_A61_len = 15;

const unsigned long long solution_an16_A61[24] = {
    65536ull, 0ull, 0ull, 1606966ull, 8059514ull, 27430662ull, 76516894ull, 175266760ull, 328771580ull,
    510287490ull, 662075682ull, 721640166ull, 661335744ull, 508583906ull, 326705966ull, 174110768ull,
    76242514ull, 26993534ull, 7476282ull, 1531440ull, 234736ull, 31156ull, 0ull, 0ull,
};
const unsigned solution_an16_A61_len = 24;

const unsigned long long solution_an24_A61[32] = {
    16777216ull, 0ull, 0ull, 1149033026ull, 7730342814ull, 37886151670ull, 157227382818ull,
    540172017744ull, 1548938717128ull, 3775178858076ull, 7912978808738ull, 14366801387648ull,
    22718650779984ull, 31423264112014ull, 38125165878518ull, 40640411783310ull, 38080772766468ull,
    31348111833932ull, 22633509878072ull, 14291703266250ull, 7859208391852ull, 3742071413564ull,
    1530719780276ull, 532399828822ull, 155292628118ull, 37269877988ull, 7160570684ull, 1060815102ull,
    115652770ull, 7976054ull, 0ull, 0ull,
};
const unsigned solution_an24_A61_len = 32;

const unsigned long long solution_an24_A641[35] = {
    16777216ull, 0ull, 10076802ull, 153434946ull, 1150452598ull, 6049205330ull, 28899516102ull,
    115368656122ull, 378189081646ull, 1058452761638ull, 2573467489292ull, 5466591303774ull,
    10201679305232ull, 16811402697690ull, 24559836981782ull, 31900986742874ull, 36914263923702ull,
    38095535515896ull, 35074592513254ull, 28799084410110ull, 21063110529820ull, 13694798923552ull,
    7892709334408ull, 4016481691866ull, 1795726693726ull, 700858354732ull, 236678653462ull,
    68351026422ull, 16618883926ull, 3276574070ull, 529150272ull, 92569066ull, 10129120ull, 3350208ull,
    0ull,
};
const unsigned solution_an24_A641_len = 35;

const unsigned long long solution_an16_A641[26] = {
    65536ull, 0ull, 0ull, 262054ull, 1471362ull, 5233458ull, 18030526ull, 51777972ull, 119084336ull,
    228590406ull, 374383580ull, 526096540ull, 636024742ull, 662561746ull, 594903286ull, 460477338ull,
    306623886ull, 174747162ull, 84584460ull, 34452696ull, 11614808ull, 3148728ull, 651320ull,
    135498ull, 45856ull, 0ull,
};
const unsigned solution_an16_A641_len = 26;

const unsigned long long solution_an8_A641[19] = {
    256ull, 0ull, 0ull, 596ull, 1302ull, 2316ull, 4624ull, 8218ull, 11060ull, 12318ull, 11024ull,
    7378ull, 3958ull, 1792ull, 552ull, 118ull, 24ull, 0ull, 0ull,
};
const unsigned solution_an8_A641_len = 19;

}  // namespace tables
}  // namespace cr
