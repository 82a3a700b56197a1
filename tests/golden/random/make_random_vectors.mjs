// Writes tests/golden/random/random_vectors.json: input/output vectors of the
// reference's counter-hash noise (shared/random.js: uniformAt and
// gaussianTriple), computed by the reference's own module. Run in the build
// container, where the reference is mounted:
//   node tests/golden/random/make_random_vectors.mjs /root/reference
// The GPU tests and the oracle's pins read only the JSON it writes.
import { writeFileSync } from "fs";
import { dirname, join } from "path";
import { fileURLToPath } from "url";

const root = process.argv[2] || "/root/reference";
const here = dirname(fileURLToPath(import.meta.url));

async function main() {
  const random = await import(join(root, "shared", "random.js"));

  const uniforms = [];
  const seeds = [0, 1, 99, 101, 202, 20261106, 0xffffffff];
  for (const seed of seeds) {
    for (const a of [0, 1, 7, 1189, 65535, 0x7fffffff]) {
      for (const b of [0, 3, 41, 0x5eed, 1000000]) {
        for (let c = 0; c < 4; c += 1) uniforms.push([seed, a, b, c, random.uniformAt(seed, a, b, c)]);
      }
    }
  }
  const triples = [];
  const out = new Float64Array(3);
  for (const [seed, step] of [[20261106, 41], [99, 3], [202, 0x5eed]]) {
    for (let atom = 0; atom < 64; atom += 1) {
      random.gaussianTriple(out, seed, atom, step);
      triples.push([seed, atom, step, out[0], out[1], out[2]]);
    }
  }
  writeFileSync(join(here, "random_vectors.json"),
    JSON.stringify({ source: "shared/random.js (uniformAt, gaussianTriple)", uniforms, triples }));
  console.log(`wrote ${uniforms.length} uniforms and ${triples.length} triples`);
}

main().catch((error) => {
  console.error(error);
  process.exit(1);
});
