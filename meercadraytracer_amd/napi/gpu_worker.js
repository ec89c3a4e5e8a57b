'use strict';
// A WorkerLike (src/executors.ts:46-51) whose postMessage drives libmcr.so through mcr.node, for
// RaytracingRenderer's workerFactory (src/RaytracingRenderer.ts:117-125):
//   new RaytracingRenderer({ workerFactory: () => createGpuWorker(), workers: 1 })
// Messages keep the shapes of src/protocol.ts:29-49 and are handled in arrival order, as the
// reference's tile handler handles them (createTileHandler, protocol.ts:53-110): geometry and
// shading replace what the worker holds, textures add pixels by id, and each tile message gets
// a 'tile' response with the premultiplied RGBA floats or an 'error' response with the message
// text. Plain ES2019, so it runs on the reference's Node and on older ones.
const path = require('path');
const mcr = require(path.join(__dirname, 'mcr.node'));

function createGpuWorker() {
  const listeners = { message: [], error: [] };
  const scene = mcr.sceneCreate();
  const pixels = new Map();  // texture id -> pixels (protocol.ts:68-70)
  let pendingShading = null; // the last 'shading' message, joined to its pixels at the next tile
  let chain = Promise.resolve();
  let closed = false;

  const emit = (type, event) => listeners[type].forEach(fn => fn(event));
  const reply = data => emit('message', { data });

  const handle = async msg => {
    switch (msg.type) {
      case 'geometry':
        mcr.sceneSetGeometry(scene, msg.geometry, msg.version);
        return;
      case 'textures':
        for (const item of msg.items) pixels.set(item.id, item.data);
        return;
      case 'shading':
        pendingShading = msg;
        return;
      case 'tile':
        try {
          if (pendingShading) {
            mcr.sceneSetShading(scene, pendingShading.shading, pixels, pendingShading.version);
            pendingShading = null;
          }
          const data = await mcr.renderTileAsync(scene, msg);
          reply({ type: 'tile', job: msg.job, data });
        } catch (error) {
          reply({ type: 'error', job: msg.job, message: error.message });
        }
        return;
      default:
        return;
    }
  };

  return {
    postMessage(msg) {
      if (closed) return;
      chain = chain.then(() => handle(msg)).catch(error => emit('error', { message: error.message }));
    },
    addEventListener(type, fn) {
      if (listeners[type]) listeners[type].push(fn);
    },
    terminate() {
      closed = true;
      chain = chain.then(() => mcr.sceneRelease(scene));
    },
  };
}

module.exports = { createGpuWorker, mcr };
